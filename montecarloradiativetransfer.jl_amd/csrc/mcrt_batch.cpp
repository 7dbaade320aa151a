// mcrt_batch.cpp — many solves in one persistent launch per precision (the C
// ABI mcrt_batch_*).  Each solve is cut into chunks (ray ranges) by a rule
// that depends on the solve alone; the chunks of all solves form one queue,
// longest estimated work first, which the workgroups of batch_kernel drain.
// The launch sequence (clear, transport, reduce) is captured into a HIP graph
// on the first launch and replayed on later ones.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "../../include/mcrt.h"
#include "mcrt_coop.h"
#include "mcrt_internal.h"
#include "mcrt_kernels.h"

using mcrt::fail;

struct mcrt_batch {
  mcrt_ctx* ctx = nullptr;
  std::vector<void*> allocs;
  size_t bytes = 0;
  std::vector<mcrt::SolveDev> solves;
  struct Launch {
    bool mixed = false;
    mcrt::BatchChunk* dchunks = nullptr;
    int nchunks = 0;
    size_t lds = 0;
    unsigned int* dnext = nullptr;
    void* slot_mem = nullptr;
  };
  std::vector<Launch> launches;
  int blocks = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool launched = false;

  ~mcrt_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (void* p : allocs) (void)hipFree(p);
    for (const mcrt::SolveDev& sd : solves) mcrt::table_cache_release(ctx->tables, sd.table_refs);
  }
  void* alloc(size_t n, hipError_t* e) {
    void* p = nullptr;
    *e = hipMalloc(&p, n ? n : 8);
    if (*e != hipSuccess) return nullptr;
    allocs.push_back(p);
    bytes += n;
    return p;
  }
};

#define HIP_OK(ctx, expr)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail((ctx), e_ == hipErrorOutOfMemory ? MCRT_ERR_NOMEM : MCRT_ERR_HIP, "%s: %s", \
                  #expr, hipGetErrorString(e_));                                              \
  } while (0)

extern "C" int mcrt_batch_create(mcrt_ctx* ctx, int32_t n, const mcrt_request* reqs,
                                 mcrt_batch** out) {
  if (!ctx) return fail(nullptr, MCRT_ERR_INVALID, "context is NULL");
  if (!out || !reqs || n < 1) return fail(ctx, MCRT_ERR_INVALID, "batch_create: no requests");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    int rc;
    if ((rc = mcrt::validate_config(ctx, reqs[i].cfg))) return rc;
    if ((rc = mcrt::validate_geometry(ctx, reqs[i].geom))) return rc;
    if (reqs[i].precision != MCRT_PRECISION_FP64 && reqs[i].precision != MCRT_PRECISION_MIXED)
      return fail(ctx, MCRT_ERR_UNSUPPORTED, "request %d: batches run the fp64 and mixed forms", i);
    if (reqs[i].chunk_rays < 0) return fail(ctx, MCRT_ERR_INVALID, "request %d: chunk_rays < 0", i);
    if (reqs[i].slots < 0 || reqs[i].slots % 64)
      return fail(ctx, MCRT_ERR_INVALID, "request %d: slots must be 0 or a positive multiple of 64", i);
  }
  HIP_OK(ctx, hipSetDevice(ctx->device));
  auto* B = new mcrt_batch();
  B->ctx = ctx;
  B->solves.resize(n);
  auto alloc = [B](size_t bytes, hipError_t* e) { return B->alloc(bytes, e); };
  for (int i = 0; i < n; ++i) {
    const mcrt_request& q = reqs[i];
    int rc = mcrt::solve_create(B->solves[i], ctx->lds_per_block, q.cfg, q.geom, q.I0, q.seed,
                                q.ray_begin, q.instance, q.precision == MCRT_PRECISION_MIXED,
                                q.chunk_rays, mcrt::ChunkRule::Throughput, alloc, &ctx->err,
                                ctx->tables, q.slots, q.stage_s);
    if (rc) {
      std::string m = ctx->err;
      delete B;
      return fail(ctx, rc, "request %d: %s", i, m.c_str());
    }
  }
  B->blocks = std::max(1, ctx->cu_count);
  hipError_t e;
  for (int pass = 0; pass < 2; ++pass) {
    const bool mixed = pass == 0;
    // the queue: chunks of every solve of this precision, longest work first
    struct Item { double cost; int solve; int64_t k; };
    std::vector<Item> items;
    size_t lds = 0;
    for (int i = 0; i < n; ++i) {
      const mcrt::SolveDev& sd = B->solves[i];
      if (sd.mixed != mixed) continue;
      lds = std::max(lds, sd.lds_bytes);
      for (int64_t k = 0; k < sd.nchunks; ++k) items.push_back({sd.cost, i, k});
    }
    if (items.empty()) continue;
    std::stable_sort(items.begin(), items.end(),
                     [](const Item& a, const Item& b) { return a.cost > b.cost; });
    if (items.size() > (size_t)0x7fffffff) {
      delete B;
      return fail(ctx, MCRT_ERR_INVALID, "batch: too many chunks");
    }
    std::vector<mcrt::BatchChunk> h(items.size());
    for (size_t j = 0; j < items.size(); ++j) mcrt::solve_chunk(B->solves[items[j].solve], items[j].k, &h[j]);
    mcrt_batch::Launch L;
    L.mixed = mixed;
    L.nchunks = (int)h.size();
    L.lds = lds;
    L.dchunks = (mcrt::BatchChunk*)B->alloc(sizeof(mcrt::BatchChunk) * h.size(), &e);
    L.dnext = (unsigned int*)B->alloc(sizeof(unsigned int), &e);
    L.slot_mem = B->alloc(mcrt::batch_slot_bytes(mixed, B->blocks), &e);
    if (!L.dchunks || !L.dnext || !L.slot_mem) {
      delete B;
      return fail(ctx, MCRT_ERR_NOMEM, "batch: allocation failed: %s", hipGetErrorString(e));
    }
    e = hipMemcpy(L.dchunks, h.data(), sizeof(mcrt::BatchChunk) * h.size(), hipMemcpyHostToDevice);
    // the attribute is per kernel and process-wide: set it to the most a batch
    // kernel can use, so that no solve's launch depends on another's setting
    if (e == hipSuccess) e = mcrt::batch_set_lds(mixed, ctx->lds_per_block - 16);
    if (e != hipSuccess) {
      delete B;
      return fail(ctx, MCRT_ERR_HIP, "batch: set-up failed: %s", hipGetErrorString(e));
    }
    B->launches.push_back(L);
  }
  e = hipEventCreate(&B->ev0);
  if (e == hipSuccess) e = hipEventCreate(&B->ev1);
  if (e != hipSuccess) {
    delete B;
    return fail(ctx, MCRT_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e));
  }
  *out = B;
  return MCRT_OK;
}

namespace {
// The whole launch sequence on stream s (captured once into the graph).
hipError_t enqueue(mcrt_batch* B, hipStream_t s) {
  hipError_t e = hipSuccess;
  for (const mcrt::SolveDev& sd : B->solves)
    if ((e = mcrt::solve_prepare(sd, s)) != hipSuccess) return e;
  for (const mcrt_batch::Launch& L : B->launches)
    if ((e = mcrt::batch_launch(L.mixed, L.dchunks, L.nchunks, L.dnext, L.slot_mem, B->blocks,
                                L.lds, s)) != hipSuccess)
      return e;
  for (const mcrt::SolveDev& sd : B->solves)
    if ((e = mcrt::solve_finish(sd, s)) != hipSuccess) return e;
  return e;
}
}  // namespace

extern "C" int mcrt_batch_launch(mcrt_batch* B) {
  if (!B) return fail(nullptr, MCRT_ERR_INVALID, "batch is NULL");
  mcrt_ctx* ctx = B->ctx;
  HIP_OK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  if (!B->exec) {
    HIP_OK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    hipError_t e = enqueue(B, s);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(ctx, MCRT_ERR_HIP, "batch: graph capture failed: %s",
                  hipGetErrorString(e != hipSuccess ? e : e2));
    }
    B->graph = g;
    HIP_OK(ctx, hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
  }
  HIP_OK(ctx, hipEventRecord(B->ev0, s));
  HIP_OK(ctx, hipGraphLaunch(B->exec, s));
  HIP_OK(ctx, hipEventRecord(B->ev1, s));
  B->launched = true;
  return MCRT_OK;
}

extern "C" int mcrt_batch_wait(mcrt_batch* B) {
  if (!B) return fail(nullptr, MCRT_ERR_INVALID, "batch is NULL");
  if (!B->launched) return fail(B->ctx, MCRT_ERR_NOT_READY, "batch was not launched");
  HIP_OK(B->ctx, hipSetDevice(B->ctx->device));
  HIP_OK(B->ctx, hipStreamSynchronize(B->ctx->stream));
  return MCRT_OK;
}

extern "C" int mcrt_batch_read(mcrt_batch* B, int32_t i, double A[5], double* Irt, double* Icb,
                               double A_sigma[5], double* Irt_sigma, double* Icb_sigma,
                               mcrt_stats* st) {
  if (!B) return fail(nullptr, MCRT_ERR_INVALID, "batch is NULL");
  mcrt_ctx* ctx = B->ctx;
  if (i < 0 || i >= (int)B->solves.size()) return fail(ctx, MCRT_ERR_INVALID, "batch_read: index %d", i);
  int rc = mcrt_batch_wait(B);
  if (rc) return rc;
  const mcrt::SolveDev& sd = B->solves[i];
  std::vector<double> v((size_t)sd.nout), sg((size_t)sd.nout);
  HIP_OK(ctx, hipMemcpy(v.data(), sd.out, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
  HIP_OK(ctx, hipMemcpy(sg.data(), sd.sig, sizeof(double) * sg.size(), hipMemcpyDeviceToHost));
  const double N = (double)sd.N;
  if (A) for (int c = 0; c < 5; ++c) A[c] = v[c] / N;
  if (Irt) for (int64_t c = 0; c < sd.nrt; ++c) Irt[c] = v[5 + c] / N;
  if (Icb) for (int64_t c = 0; c < sd.ncb; ++c) Icb[c] = v[5 + sd.nrt + c] / N;
  if (A_sigma) for (int c = 0; c < 5; ++c) A_sigma[c] = sg[c];
  if (Irt_sigma) for (int64_t c = 0; c < sd.nrt; ++c) Irt_sigma[c] = sg[5 + c];
  if (Icb_sigma) for (int64_t c = 0; c < sd.ncb; ++c) Icb_sigma[c] = sg[5 + sd.nrt + c];
  if (st) {
    unsigned long long c[mcrt::kCntN];
    HIP_OK(ctx, hipMemcpy(c, sd.counters, sizeof c, hipMemcpyDeviceToHost));
    st->rays = c[mcrt::kCntRays];
    st->events = c[mcrt::kCntEvents];
    st->rt_evals = c[mcrt::kCntRt];
    st->cb_evals = c[mcrt::kCntCb];
    st->cb_evals_single = c[mcrt::kCntCb1];
    st->rejections = c[mcrt::kCntRej];
    st->nonfinite = c[mcrt::kCntNonfinite];
    double ms = 0;
    mcrt_batch_kernel_ms(B, &ms);
    st->kernel_ms = ms;
    st->total_ms = ms;
    st->batches = (uint64_t)sd.nbatch;
    st->chunks = (uint64_t)sd.nchunks;
  }
  return MCRT_OK;
}

extern "C" int mcrt_batch_kernel_ms(mcrt_batch* B, double* ms) {
  if (!B || !ms) return fail(nullptr, MCRT_ERR_INVALID, "batch_kernel_ms: NULL");
  if (!B->launched) return fail(B->ctx, MCRT_ERR_NOT_READY, "batch was not launched");
  float f = 0.f;
  HIP_OK(B->ctx, hipEventElapsedTime(&f, B->ev0, B->ev1));
  *ms = f;
  return MCRT_OK;
}

extern "C" int mcrt_batch_info(mcrt_batch* B, int64_t* chunks, int64_t* device_bytes) {
  if (!B) return fail(nullptr, MCRT_ERR_INVALID, "batch is NULL");
  int64_t c = 0;
  for (const auto& L : B->launches) c += L.nchunks;
  if (chunks) *chunks = c;
  if (device_bytes) *device_bytes = (int64_t)B->bytes;
  return MCRT_OK;
}

extern "C" void mcrt_batch_destroy(mcrt_batch* B) {
  if (!B) return;
  (void)hipSetDevice(B->ctx->device);
  (void)hipStreamSynchronize(B->ctx->stream);
  delete B;
}
