// mcrt_static.hip — mcrt_simulate_static: a solve that allocates no device
// memory, for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops, harnesses that forbid allocation).
//
// Everything the solve needs on the device beyond the caller's inputs and
// outputs — the parameter block, bin constants, scatterer tables, batch slabs,
// ray-slot state, chunk queue and counters — is carved out of one file-scope
// device array, `scratch`, by a bump allocator.  The transport is the batched
// solve of mcrt_batch_* for a single request (zero-fill kernels, batch_kernel,
// batch_reduce), so the result equals a one-request batch of the same solve up
// to the order of FP64 additions.
//
// Each device holds its own copy of `scratch` at its own address, found once
// per device with hipGetSymbolAddress when a context is created on it.  Every
// call on a device shares that device's array and leaves the last call's
// contents there: what must start at zero is cleared by kernels at the start
// of each call.  Calls that overlap on one device (from two threads, two
// contexts or two processes' worth of contexts on one device) race on the
// array and are not supported.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <mutex>

#include "../../include/mcrt.h"
#include "mcrt_coop.h"
#include "mcrt_internal.h"
#include "mcrt_kernels.h"

namespace mcrt {

// 96 MiB: the sphere Mie CB workload at 180 × 48 RT and up to 276 × 8 CB bins
// needs about 75 MB (128 batch slabs of the outputs, 256 workgroups of ray-slot
// state, tables and bin constants).
constexpr size_t kScratchBytes = size_t(96) << 20;
constexpr int kMaxDevices = 64;

namespace {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
hipError_t g_status[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

}  // namespace

// Device address of `scratch` on `device`, which must be current.  Called by
// mcrt_context_create; the status is kept so that mcrt_simulate_static can
// report why the array is unavailable.
hipError_t static_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_done[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_status[device] = e;
    g_ptr[device] = e == hipSuccess ? p : nullptr;
    g_done[device] = true;
  }
  return g_status[device];
}

}  // namespace mcrt

extern "C" int mcrt_simulate_static(mcrt_ctx* ctx, const mcrt_config* cfg, const mcrt_geometry* geom,
                                    const double I0[4], uint64_t seed, int32_t precision, double A[5],
                                    double* Irt, double* Icb, mcrt_stats* stats) {
  using mcrt::fail;
  const auto t0 = std::chrono::steady_clock::now();
  if (!ctx) return fail(nullptr, MCRT_ERR_INVALID, "context is NULL");
  if (!I0 || !A || !Irt) return fail(ctx, MCRT_ERR_INVALID, "simulate_static: NULL argument");
  int rc;
  if ((rc = mcrt::validate_config(ctx, cfg))) return rc;
  if ((rc = mcrt::validate_geometry(ctx, geom))) return rc;
  if (precision != MCRT_PRECISION_FP64 && precision != MCRT_PRECISION_MIXED)
    return fail(ctx, MCRT_ERR_UNSUPPORTED,
                "simulate_static runs the fp64 and mixed forms (the plain form keeps a path "
                "history of maximum_scattering_times per thread)");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return fail(ctx, MCRT_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
  e = mcrt::static_scratch_init(ctx->device);
  if (e != hipSuccess || !mcrt::g_ptr[ctx->device])
    return fail(ctx, MCRT_ERR_HIP, "simulate_static: no scratch array on device %d: %s", ctx->device,
                hipGetErrorString(e));
  unsigned char* const base = (unsigned char*)mcrt::g_ptr[ctx->device];
  // bump allocation over the scratch array (never freed: the next call starts over)
  size_t off = 0, need = 0;
  const mcrt::DevAlloc bump = [&](size_t n, hipError_t* err) -> void* {
    const size_t a = (off + 255) & ~size_t(255);
    need = std::max(need, a + n);
    if (a + n > mcrt::kScratchBytes) {
      *err = hipErrorOutOfMemory;
      off = a + n;
      return nullptr;
    }
    off = a + n;
    *err = hipSuccess;
    return base + a;
  };
  const bool mixed = precision == MCRT_PRECISION_MIXED;
  mcrt::SolveDev sd;
  rc = mcrt::solve_create(sd, ctx->lds_per_block, cfg, geom, I0, seed, 0, 0, mixed, 0,
                          mcrt::ChunkRule::Throughput, bump, &ctx->err, nullptr);
  const int blocks = std::max(1, ctx->cu_count);
  mcrt::BatchChunk* dchunks = nullptr;
  unsigned int* dnext = nullptr;
  void* slot_mem = nullptr;
  if (rc == MCRT_OK) {
    dchunks = (mcrt::BatchChunk*)bump(sizeof(mcrt::BatchChunk) * (size_t)sd.nchunks, &e);
    dnext = (unsigned int*)bump(sizeof(unsigned int), &e);
    slot_mem = bump(mcrt::batch_slot_bytes(mixed, blocks), &e);
    if (!dchunks || !dnext || !slot_mem) rc = MCRT_ERR_NOMEM;
  }
  if (rc == MCRT_ERR_NOMEM)
    return fail(ctx, MCRT_ERR_NOMEM,
                "simulate_static: the solve needs %zu bytes of device scratch, more than the "
                "fixed %zu (fewer bins, or mcrt_simulate)",
                need, mcrt::kScratchBytes);
  if (rc) return fail(ctx, rc, "simulate_static: %s", ctx->err.c_str());
  if (sd.nchunks > 0x7fffffff) return fail(ctx, MCRT_ERR_INVALID, "simulate_static: too many chunks");
  // the chunk queue, staged through a small buffer on the stack
  hipStream_t s = ctx->stream;
  {
    mcrt::BatchChunk buf[256];
    for (int64_t k0 = 0; k0 < sd.nchunks && e == hipSuccess; k0 += 256) {
      const int64_t m = std::min<int64_t>(256, sd.nchunks - k0);
      for (int64_t j = 0; j < m; ++j) mcrt::solve_chunk(sd, k0 + j, &buf[j]);
      e = hipMemcpy(dchunks + k0, buf, sizeof(mcrt::BatchChunk) * (size_t)m, hipMemcpyHostToDevice);
    }
  }
  if (e == hipSuccess) e = mcrt::batch_set_lds(mixed, ctx->lds_per_block - 16);  // as every batch path
  if (e == hipSuccess) e = mcrt::solve_prepare(sd, s);  // slabs and counters cleared by kernels
  if (e == hipSuccess)
    e = mcrt::batch_launch(mixed, dchunks, (int)sd.nchunks, dnext, slot_mem, blocks, sd.lds_bytes, s);
  if (e == hipSuccess) e = mcrt::solve_finish(sd, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  const auto t1 = std::chrono::steady_clock::now();
  // read-back straight into the caller's arrays, normalised by N on the host
  if (e == hipSuccess) e = hipMemcpy(A, sd.out, sizeof(double) * 5, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(Irt, sd.out + 5, sizeof(double) * (size_t)sd.nrt, hipMemcpyDeviceToHost);
  if (e == hipSuccess && Icb && sd.ncb > 0)  // Icb may be NULL, as for mcrt_plan_read
    e = hipMemcpy(Icb, sd.out + 5 + sd.nrt, sizeof(double) * (size_t)sd.ncb, hipMemcpyDeviceToHost);
  unsigned long long c[mcrt::kCntN] = {};
  if (e == hipSuccess && stats) e = hipMemcpy(c, sd.counters, sizeof c, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail(ctx, MCRT_ERR_HIP, "simulate_static: %s", hipGetErrorString(e));
  const double N = (double)sd.N;
  for (int i = 0; i < 5; ++i) A[i] /= N;
  for (int64_t i = 0; i < sd.nrt; ++i) Irt[i] /= N;
  if (Icb) for (int64_t i = 0; i < sd.ncb; ++i) Icb[i] /= N;
  if (stats) {
    stats->rays = c[mcrt::kCntRays];
    stats->events = c[mcrt::kCntEvents];
    stats->rt_evals = c[mcrt::kCntRt];
    stats->cb_evals = c[mcrt::kCntCb];
    stats->cb_evals_single = c[mcrt::kCntCb1];
    stats->rejections = c[mcrt::kCntRej];
    stats->nonfinite = c[mcrt::kCntNonfinite];
    stats->kernel_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    stats->batches = (uint64_t)sd.nbatch;
    stats->chunks = (uint64_t)sd.nchunks;
  }
  return MCRT_OK;
}
