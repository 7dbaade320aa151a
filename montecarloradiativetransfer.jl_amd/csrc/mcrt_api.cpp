// mcrt_api.cpp — the C ABI of libmcrt_hip.so: contexts, plans, validation,
// table upload, launch and read-back.  The solve itself is
// simulate(cfg, g, I₀; use_gpu=true), src/MonteCarloRadiativeTransfer.jl:26-85.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mcrt.h"
#include "mcrt_coop.h"
#include "mcrt_internal.h"
#include "mcrt_kernels.h"

namespace {

thread_local std::string g_thread_err;
constexpr double kPi = 3.141592653589793;
// Device bytes of scatterer tables a context keeps for later solves once no
// plan or handle uses them (least recently used go first beyond this).
constexpr size_t kContextTableBytes = size_t(1) << 30;

int validate_scatterer(mcrt_ctx* ctx, const mcrt_scatterer* s);

// A device allocation that frees itself.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t n) {
    release();
    if (n == 0) n = 8;
    bytes = n;
    return hipMalloc(&p, n);
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace

struct mcrt_plan {
  mcrt_ctx* ctx = nullptr;
  int precision = 0;
  int64_t N = 0;
  int64_t nrt = 0, ncb = 0;
  int geo_kind = 0;
  // device memory
  std::vector<DevBuf*> bufs;
  double* dA = nullptr;
  double* dIrt = nullptr;
  double* dIcb = nullptr;
  unsigned long long* dcnt = nullptr;
  // plain form
  mcrt::PlainParams plain{};
  mcrt::PlainParams* dplain = nullptr;
  int plain_blocks = 0;
  // cooperative forms
  mcrt::CoopPlan coop{};
  double* dsig = nullptr;  // batch-means standard errors (allocated on first use)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool launched = false;
  std::chrono::steady_clock::time_point t_launch;
  double total_ms = 0.0;

  ~mcrt_plan() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (DevBuf* b : bufs) delete b;
    if (ctx) mcrt::table_cache_release(ctx->tables, coop.table_refs);
  }
  // Allocate n bytes of device memory owned by the plan.
  void* alloc(size_t n, hipError_t* e) {
    DevBuf* b = new DevBuf();
    *e = b->alloc(n);
    if (*e != hipSuccess) { delete b; return nullptr; }
    bufs.push_back(b);
    return b->p;
  }
};

namespace mcrt {

int fail(mcrt_ctx* ctx, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  g_thread_err = buf;
  return code;
}

}  // namespace mcrt

using mcrt::fail;

#define HIP_TRY(ctx, expr)                                                              \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail((ctx), e_ == hipErrorOutOfMemory ? MCRT_ERR_NOMEM : MCRT_ERR_HIP,      \
                  "%s: %s", #expr, hipGetErrorString(e_));                              \
  } while (0)

extern "C" const char* mcrt_version(void) { return "mcrt 0.2.0 (gfx950, ABI 2)"; }

extern "C" const char* mcrt_last_error(const mcrt_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  return g_thread_err.c_str();
}

extern "C" int mcrt_device_count(int* count) {
  if (!count) return fail(nullptr, MCRT_ERR_INVALID, "device_count: NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *count = (e == hipSuccess) ? n : 0;
  return MCRT_OK;
}

extern "C" int mcrt_context_create(int device, mcrt_ctx** out) {
  if (!out) return fail(nullptr, MCRT_ERR_INVALID, "context_create: out is NULL");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(nullptr, MCRT_ERR_NO_DEVICE, "no HIP device is visible");
  if (device < 0 || device >= n)
    return fail(nullptr, MCRT_ERR_INVALID, "device %d out of range (0..%d)", device, n - 1);
  auto* c = new mcrt_ctx();
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  hipDeviceProp_t prop;
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    int rc = fail(nullptr, MCRT_ERR_HIP, "context_create: %s", hipGetErrorString(e));
    delete c;
    return rc;
  }
  c->cu_count = prop.multiProcessorCount;
  c->lds_per_block = prop.sharedMemPerBlock;  // 160 KiB on gfx950
  // the no-allocation solve's scratch address on this device; a failure is
  // kept and reported by mcrt_simulate_static, the other calls do not need it
  (void)mcrt::static_scratch_init(device);
  c->tables = mcrt::table_cache_create(kContextTableBytes);
  *out = c;
  return MCRT_OK;
}

extern "C" void mcrt_context_destroy(mcrt_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
  }
  for (auto& kv : ctx->table_handles) mcrt::table_cache_release(ctx->tables, kv.second);
  mcrt::table_cache_destroy(ctx->tables);
  delete ctx;
}

extern "C" int mcrt_table_upload(mcrt_ctx* ctx, const mcrt_scatterer* sca, uint64_t* handle) {
  if (!ctx || !sca || !handle) return fail(ctx, MCRT_ERR_INVALID, "table_upload: NULL argument");
  *handle = 0;
  int rc = validate_scatterer(ctx, sca);
  if (rc) return rc;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::vector<const void*> refs;
  std::string err;
  rc = mcrt::table_pin(ctx->tables, sca, &refs, &err);
  if (rc) {
    mcrt::table_cache_release(ctx->tables, refs);
    return fail(ctx, rc, "table_upload: %s", err.c_str());
  }
  std::lock_guard<std::mutex> g(ctx->table_mu);
  *handle = ctx->next_table++;
  ctx->table_handles[*handle] = std::move(refs);
  return MCRT_OK;
}

extern "C" int mcrt_table_release(mcrt_ctx* ctx, uint64_t handle) {
  if (!ctx) return fail(nullptr, MCRT_ERR_INVALID, "table_release: context is NULL");
  std::lock_guard<std::mutex> g(ctx->table_mu);
  auto it = ctx->table_handles.find(handle);
  if (it == ctx->table_handles.end())
    return fail(ctx, MCRT_ERR_INVALID, "table_release: unknown handle %llu", (unsigned long long)handle);
  (void)hipSetDevice(ctx->device);
  mcrt::table_cache_release(ctx->tables, it->second);
  ctx->table_handles.erase(it);
  return MCRT_OK;
}

extern "C" int mcrt_table_cache_stats(mcrt_ctx* ctx, int64_t* bytes, uint64_t* hits, uint64_t* misses) {
  if (!ctx) return fail(nullptr, MCRT_ERR_INVALID, "table_cache_stats: context is NULL");
  size_t b = 0;
  mcrt::table_cache_stats(ctx->tables, &b, hits, misses);
  if (bytes) *bytes = (int64_t)b;
  return MCRT_OK;
}

namespace {

int validate_scatterer(mcrt_ctx* ctx, const mcrt_scatterer* s) {
  if (!s) return fail(ctx, MCRT_ERR_INVALID, "scatterer is NULL");
  if (s->kind == MCRT_SCATTERER_RAYLEIGH) return MCRT_OK;
  if (s->kind != MCRT_SCATTERER_MIE && s->kind != MCRT_SCATTERER_SPLINE)
    return fail(ctx, MCRT_ERR_INVALID, "unknown scatterer kind %d", s->kind);
  const char* name = s->kind == MCRT_SCATTERER_MIE ? "Mie" : "Spline";
  if (s->n_theta < 2 || s->random_points < 1)
    return fail(ctx, MCRT_ERR_INVALID, "%s scatterer needs Nθ ≥ 2 and random_points ≥ 1", name);
  if (!s->csrn || !s->F11 || !s->F12 || !s->F33 || !s->F34 || !s->S11 || !s->S22 ||
      (s->kind == MCRT_SCATTERER_SPLINE && (!s->F22 || !s->F44)))
    return fail(ctx, MCRT_ERR_INVALID, "%s scatterer table pointer is NULL", name);
  return MCRT_OK;
}

}  // namespace

namespace mcrt {

// validate(p::PlaneGeometry), plane.jl:52-62; validate(s::SphereGeometry), sphere.jl:35-37.
int validate_geometry(mcrt_ctx* ctx, const mcrt_geometry* g) {
  if (!g) return fail(ctx, MCRT_ERR_INVALID, "geometry is NULL");
  if (g->kind == MCRT_GEOMETRY_PLANE) {
    if (!(kPi / 2 < g->theta && g->theta <= kPi)) return fail(ctx, MCRT_ERR_INVALID, "θ must be in (π/2, π]");
    if (!(0 <= g->phi && g->phi < 2 * kPi)) return fail(ctx, MCRT_ERR_INVALID, "ϕ must be in [0, 2π)");
    if (g->n_layers < 1 || !g->layers) return fail(ctx, MCRT_ERR_INVALID, "There needs to be at least one layer");
    for (int i = 0; i < g->n_layers; ++i)
      if (i != g->n_layers - 1 && g->layers[i].d < 0)
        return fail(ctx, MCRT_ERR_INVALID, "Only the last layer can be infinite");
    if (g->n_layers > mcrt::kMaxLayers)
      return fail(ctx, MCRT_ERR_UNSUPPORTED, "at most %d layers are supported", mcrt::kMaxLayers);
    for (int i = 0; i < g->n_layers; ++i) {
      int rc = validate_scatterer(ctx, g->layers[i].scatterer);
      if (rc) return rc;
    }
  } else if (g->kind == MCRT_GEOMETRY_SPHERE) {
    if (!(g->R > 0)) return fail(ctx, MCRT_ERR_INVALID, "Radius must be positive");
    int rc = validate_scatterer(ctx, g->scatterer);
    if (rc) return rc;
  } else {
    return fail(ctx, MCRT_ERR_INVALID, "unknown geometry kind %d", g->kind);
  }
  return MCRT_OK;
}

int validate_config(mcrt_ctx* ctx, const mcrt_config* c) {
  if (!c) return fail(ctx, MCRT_ERR_INVALID, "config is NULL");
  if (c->number_of_rays < 1) return fail(ctx, MCRT_ERR_INVALID, "number_of_rays must be ≥ 1");
  if (c->maximum_scattering_times < 0) return fail(ctx, MCRT_ERR_INVALID, "maximum_scattering_times must be ≥ 0");
  if (c->n_theta < 1 || c->n_phi < 1) return fail(ctx, MCRT_ERR_INVALID, "Nθ and Nϕ must be ≥ 1");
  if (!c->cos_theta || !c->sin_theta || !c->inorm || !c->cos_phi || !c->sin_phi)
    return fail(ctx, MCRT_ERR_INVALID, "RT angle vectors must be given");
  if (c->n_theta_b < 0 || c->n_phi_b < 0) return fail(ctx, MCRT_ERR_INVALID, "Nθb and Nϕb must be ≥ 0");
  if (c->compute_cb && (c->n_theta_b < 1 || c->n_phi_b < 1 || !c->cos_theta_b || !c->sin_theta_b ||
                        !c->cos_phi_b || !c->sin_phi_b))
    return fail(ctx, MCRT_ERR_INVALID, "compute_cb needs the CB angle vectors");
  return MCRT_OK;
}

}  // namespace mcrt

namespace {

template <typename T>
T* upload(mcrt_plan* P, const T* host, size_t n, int* rc) {
  hipError_t e;
  T* d = static_cast<T*>(P->alloc(sizeof(T) * (n ? n : 1), &e));
  if (!d) { *rc = fail(P->ctx, e == hipErrorOutOfMemory ? MCRT_ERR_NOMEM : MCRT_ERR_HIP, "hipMalloc: %s", hipGetErrorString(e)); return nullptr; }
  if (n) {
    e = hipMemcpy(d, host, sizeof(T) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) { *rc = fail(P->ctx, MCRT_ERR_HIP, "hipMemcpy: %s", hipGetErrorString(e)); return nullptr; }
  }
  *rc = MCRT_OK;
  return d;
}

// Distinct scatterers of a geometry, in layer order.
std::vector<const mcrt_scatterer*> scatterers_of(const mcrt_geometry* g, int* layer_slot) {
  std::vector<const mcrt_scatterer*> v;
  int nl = g->kind == MCRT_GEOMETRY_PLANE ? g->n_layers : 1;
  for (int i = 0; i < nl; ++i) {
    const mcrt_scatterer* s = g->kind == MCRT_GEOMETRY_PLANE ? g->layers[i].scatterer : g->scatterer;
    int slot = -1;
    for (size_t j = 0; j < v.size(); ++j)
      if (v[j] == s) slot = (int)j;
    if (slot < 0) { slot = (int)v.size(); v.push_back(s); }
    layer_slot[i] = slot;
  }
  return v;
}

// Interleave a scatterer's tables: F per node (F11, F12, F33, F34) for Mie; for
// Spline that block over N + 1 nodes (the last zero), then (F22, F44) per node
// over N + 1 nodes (ScaView); S per node (S11 re, S11 im, S22 re, S22 im).
template <typename T>
void interleave(const mcrt_scatterer* s, std::vector<T>& F, std::vector<T>& S, std::vector<T>& csrn) {
  const int N = s->n_theta;
  const bool spl = s->kind == MCRT_SCATTERER_SPLINE;
  F.assign(spl ? (size_t)mcrt::f_stride(s->kind) * (N + 1) : 4 * (size_t)N, T(0));
  S.resize(4 * (size_t)N);
  csrn.resize((size_t)s->random_points + 1);
  for (int i = 0; i < N; ++i) {
    T* f = &F[4 * (size_t)i];
    f[0] = (T)s->F11[i]; f[1] = (T)s->F12[i]; f[2] = (T)s->F33[i]; f[3] = (T)s->F34[i];
    if (spl) {
      T* b = &F[(size_t)mcrt::spl_b(N) + 2 * (size_t)i];
      b[0] = (T)s->F22[i]; b[1] = (T)s->F44[i];
    }
    S[4 * i + 0] = (T)s->S11[2 * i]; S[4 * i + 1] = (T)s->S11[2 * i + 1];
    S[4 * i + 2] = (T)s->S22[2 * i]; S[4 * i + 3] = (T)s->S22[2 * i + 1];
  }
  for (int i = 0; i <= s->random_points; ++i) csrn[i] = (T)s->csrn[i];
}

}  // namespace

namespace mcrt {

// Incident frame and k₀ (simulate, MonteCarloRadiativeTransfer.jl:27-34).
void incident(const mcrt_geometry* g, double k0[3], double f0[4]) {
  double th = g->kind == MCRT_GEOMETRY_PLANE ? g->theta : kPi;
  double ph = g->kind == MCRT_GEOMETRY_PLANE ? g->phi : 0.0;
  f0[0] = std::cos(th); f0[1] = std::sin(th); f0[2] = std::cos(ph); f0[3] = std::sin(ph);
  // ray_to_norm((0, 0, 1), cosθ₀, sinθ₀, cosϕ₀, sinϕ₀)
  k0[0] = 0.0 * f0[0] * f0[2] - 0.0 * f0[3] + 1.0 * f0[1] * f0[2];
  k0[1] = 0.0 * f0[0] * f0[3] + 0.0 * f0[2] + 1.0 * f0[1] * f0[3];
  k0[2] = -0.0 * f0[1] + 1.0 * f0[0];
}

GeoDev make_geo(const mcrt_geometry* g, int* layer_slot) {
  GeoDev G{};
  G.kind = g->kind == MCRT_GEOMETRY_PLANE ? 0 : 1;
  if (g->kind == MCRT_GEOMETRY_PLANE) {
    G.nl = g->n_layers;
    for (int i = 0; i < G.nl; ++i) {
      G.tau[i] = g->layers[i].tau;
      G.omega[i] = g->layers[i].omega;
      G.isinf[i] = g->layers[i].d < 0;
      G.sca[i] = layer_slot[i];
    }
    G.kl1 = g->layers[0].scatterer->kl;
  } else {
    G.nl = 1;
    G.tau[0] = g->tau;
    G.omega[0] = g->omega;
    G.sca[0] = layer_slot[0];
    G.kl1 = g->scatterer->kl;
  }
  return G;
}

}  // namespace mcrt

namespace {

int upload_bins(mcrt_plan* P, const mcrt_config* c, mcrt::BinsDev* B) {
  int rc = 0;
  B->Nt = (int)c->n_theta; B->Np = (int)c->n_phi;
  B->Ntb = c->compute_cb ? (int)c->n_theta_b : 0;
  B->Npb = c->compute_cb ? (int)c->n_phi_b : 0;
  if (!(B->ct = upload(P, c->cos_theta, c->n_theta, &rc))) return rc;
  if (!(B->st = upload(P, c->sin_theta, c->n_theta, &rc))) return rc;
  if (!(B->inorm = upload(P, c->inorm, c->n_theta, &rc))) return rc;
  if (!(B->cp = upload(P, c->cos_phi, c->n_phi, &rc))) return rc;
  if (!(B->sp = upload(P, c->sin_phi, c->n_phi, &rc))) return rc;
  size_t ntb = B->Ntb, npb = B->Npb;
  if (!(B->ctb = upload(P, c->cos_theta_b, ntb, &rc))) return rc;
  if (!(B->stb = upload(P, c->sin_theta_b, ntb, &rc))) return rc;
  if (!(B->cpb = upload(P, c->cos_phi_b, npb, &rc))) return rc;
  if (!(B->spb = upload(P, c->sin_phi_b, npb, &rc))) return rc;
  return MCRT_OK;
}

int create_plain(mcrt_plan* P, const mcrt_config* c, const mcrt_geometry* g, const double I0[4],
                 uint64_t seed, uint64_t ray0, uint32_t instance) {
  mcrt::PlainParams& p = P->plain;
  int rc = 0;
  int slot[mcrt::kMaxLayers];
  auto scas = scatterers_of(g, slot);
  if ((int)scas.size() > mcrt::kMaxSca) return fail(P->ctx, MCRT_ERR_UNSUPPORTED, "too many scatterers");
  p.nsca = (int)scas.size();
  for (size_t i = 0; i < scas.size(); ++i) {
    const mcrt_scatterer* s = scas[i];
    mcrt::ScaView<double>& v = p.sca[i];
    v.kind = s->kind == MCRT_SCATTERER_MIE ? 1 : (s->kind == MCRT_SCATTERER_SPLINE ? 2 : 0);
    v.Nt = s->n_theta; v.rp = s->random_points;
    v.scale = 0; v.csrn = nullptr; v.F = nullptr; v.S = nullptr;
    if (v.kind != 0) {
      std::vector<double> F, S, cs;
      interleave(s, F, S, cs);
      if (!(v.F = upload(P, F.data(), F.size(), &rc))) return rc;
      if (!(v.S = upload(P, S.data(), S.size(), &rc))) return rc;
      if (!(v.csrn = upload(P, cs.data(), cs.size(), &rc))) return rc;
      v.scale = (s->n_theta - 1) / kPi;
    }
  }
  p.geo = mcrt::make_geo(g, slot);
  if ((rc = upload_bins(P, c, &p.bins))) return rc;
  p.N = c->number_of_rays;
  p.maxscat = c->maximum_scattering_times;
  p.min_int = c->minimum_intensity;
  p.tau0 = c->tau0;
  p.compute_cb = c->compute_cb;
  for (int i = 0; i < 4; ++i) p.I0[i] = I0[i];
  mcrt::incident(g, p.k0, p.f0);
  p.seed = seed;
  p.ray0 = ray0;
  p.instance = instance;
  // threads = min(maximum_threads, number_of_rays), blocks of 256 (MonteCarloRadiativeTransfer.jl:39,46)
  int64_t mt = c->maximum_threads > 0 ? c->maximum_threads : 8192;
  int64_t threads = mt < c->number_of_rays ? mt : c->number_of_rays;
  P->plain_blocks = (int)((threads + 255) / 256);
  p.threads = threads;
  int64_t ms = p.maxscat > 0 ? p.maxscat : 1;
  hipError_t e;
  size_t kbytes = sizeof(double) * 3 * (size_t)ms * (size_t)threads;
  p.kpath = static_cast<double*>(P->alloc(kbytes, &e));
  if (!p.kpath) return fail(P->ctx, MCRT_ERR_NOMEM, "path history (%zu B): %s", kbytes, hipGetErrorString(e));
  p.xpath = static_cast<double*>(P->alloc(kbytes, &e));
  if (!p.xpath) return fail(P->ctx, MCRT_ERR_NOMEM, "path history (%zu B): %s", kbytes, hipGetErrorString(e));
  p.idxpath = static_cast<int8_t*>(P->alloc((size_t)ms * (size_t)threads, &e));
  if (!p.idxpath) return fail(P->ctx, MCRT_ERR_NOMEM, "path history: %s", hipGetErrorString(e));
  p.A = P->dA; p.Irt = P->dIrt; p.Icb = P->dIcb; p.counters = P->dcnt;
  if (!(P->dplain = upload(P, &p, 1, &rc))) return rc;
  return MCRT_OK;
}

}  // namespace

extern "C" int mcrt_plan_create_shaped(mcrt_ctx* ctx, const mcrt_config* cfg,
                                       const mcrt_geometry* geom, const double I0[4], uint64_t seed,
                                       int32_t precision, uint64_t ray_begin, uint32_t instance,
                                       const mcrt_launch_shape* shape_in, mcrt_plan** out) {
  const mcrt_launch_shape shape = shape_in ? *shape_in : mcrt_launch_shape{0, 0, 0, 0};
  if (!ctx) return fail(nullptr, MCRT_ERR_INVALID, "context is NULL");
  if (!out || !I0) return fail(ctx, MCRT_ERR_INVALID, "plan_create: NULL argument");
  *out = nullptr;
  int rc;
  if ((rc = mcrt::validate_config(ctx, cfg))) return rc;
  if ((rc = mcrt::validate_geometry(ctx, geom))) return rc;
  if (precision < MCRT_PRECISION_FP64_PLAIN || precision > MCRT_PRECISION_MIXED)
    return fail(ctx, MCRT_ERR_INVALID, "unknown precision %d", precision);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  auto* P = new mcrt_plan();
  P->ctx = ctx;
  P->precision = precision;
  P->N = cfg->number_of_rays;
  P->geo_kind = geom->kind;
  P->nrt = 4 * (int64_t)cfg->n_theta * cfg->n_phi;
  P->ncb = 4 * (int64_t)cfg->n_theta_b * cfg->n_phi_b;
  hipError_t e;
  P->dA = static_cast<double*>(P->alloc(sizeof(double) * 5, &e));
  if (P->dA) P->dIrt = static_cast<double*>(P->alloc(sizeof(double) * P->nrt, &e));
  if (P->dIrt) P->dIcb = static_cast<double*>(P->alloc(sizeof(double) * (P->ncb ? P->ncb : 1), &e));
  if (P->dIcb) P->dcnt = static_cast<unsigned long long*>(P->alloc(sizeof(unsigned long long) * mcrt::kCntN, &e));
  if (!P->dcnt) {
    rc = fail(ctx, MCRT_ERR_NOMEM, "output allocation: %s", hipGetErrorString(e));
    delete P;
    return rc;
  }
  if (precision == MCRT_PRECISION_FP64_PLAIN) {
    rc = create_plain(P, cfg, geom, I0, seed, ray_begin, instance);
  } else {
    rc = mcrt::coop_create(P->coop, ctx->stream, ctx->cu_count, ctx->lds_per_block, cfg, geom, I0, seed,
                           ray_begin, instance, precision == MCRT_PRECISION_MIXED, shape,
                           [P](size_t n, hipError_t* err) { return P->alloc(n, err); },
                           P->dA, P->dIrt, P->dIcb, P->dcnt, &ctx->err, ctx->tables);
    if (rc) g_thread_err = ctx->err;
  }
  if (rc == MCRT_OK) {
    e = hipEventCreate(&P->ev0);
    if (e == hipSuccess) e = hipEventCreate(&P->ev1);
    if (e != hipSuccess) rc = fail(ctx, MCRT_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e));
  }
  if (rc) {
    delete P;
    return rc;
  }
  *out = P;
  return MCRT_OK;
}

extern "C" int mcrt_plan_create_range(mcrt_ctx* ctx, const mcrt_config* cfg,
                                      const mcrt_geometry* geom, const double I0[4], uint64_t seed,
                                      int32_t precision, uint64_t ray_begin, uint32_t instance,
                                      mcrt_plan** out) {
  return mcrt_plan_create_shaped(ctx, cfg, geom, I0, seed, precision, ray_begin, instance, nullptr,
                                 out);
}

extern "C" int mcrt_plan_create(mcrt_ctx* ctx, const mcrt_config* cfg, const mcrt_geometry* geom,
                                const double I0[4], uint64_t seed, int32_t precision,
                                mcrt_plan** out) {
  return mcrt_plan_create_shaped(ctx, cfg, geom, I0, seed, precision, 0, 0, nullptr, out);
}

extern "C" int mcrt_plan_shape(mcrt_plan* P, mcrt_launch_shape* shape, int32_t* workgroups) {
  if (!P) return fail(nullptr, MCRT_ERR_INVALID, "plan is NULL");
  const bool plain = P->precision == MCRT_PRECISION_FP64_PLAIN;
  if (shape) {
    shape->slots = plain ? 0 : P->coop.slots;
    shape->workgroups_per_cu = plain ? 0 : P->coop.wg_per_cu;
    shape->stage_s = plain ? 0 : (P->coop.stage_s ? 1 : -1);
    shape->reserved = 0;
  }
  if (workgroups) *workgroups = plain ? P->plain_blocks : P->coop.blocks;
  return MCRT_OK;
}

extern "C" int mcrt_plan_launch(mcrt_plan* P) {
  if (!P) return fail(nullptr, MCRT_ERR_INVALID, "plan is NULL");
  mcrt_ctx* ctx = P->ctx;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  P->t_launch = std::chrono::steady_clock::now();
  hipStream_t s = ctx->stream;
  HIP_TRY(ctx, hipMemsetAsync(P->dA, 0, sizeof(double) * 5, s));
  HIP_TRY(ctx, hipMemsetAsync(P->dIrt, 0, sizeof(double) * P->nrt, s));
  HIP_TRY(ctx, hipMemsetAsync(P->dIcb, 0, sizeof(double) * (P->ncb ? P->ncb : 1), s));
  HIP_TRY(ctx, hipMemsetAsync(P->dcnt, 0, sizeof(unsigned long long) * mcrt::kCntN, s));
  if (P->precision != MCRT_PRECISION_FP64_PLAIN) HIP_TRY(ctx, mcrt::coop_prepare(P->coop, s));
  // ev0/ev1 bracket the transport kernel alone
  HIP_TRY(ctx, hipEventRecord(P->ev0, s));
  if (P->precision == MCRT_PRECISION_FP64_PLAIN) {
    HIP_TRY(ctx, mcrt::launch_plain(P->dplain, P->plain_blocks, 256, s));
  } else {
    HIP_TRY(ctx, mcrt::coop_launch(P->coop, s));
  }
  HIP_TRY(ctx, hipEventRecord(P->ev1, s));
  if (P->precision != MCRT_PRECISION_FP64_PLAIN) HIP_TRY(ctx, mcrt::coop_reduce(P->coop, s));
  P->launched = true;
  return MCRT_OK;
}

extern "C" int mcrt_plan_wait(mcrt_plan* P) {
  if (!P) return fail(nullptr, MCRT_ERR_INVALID, "plan is NULL");
  if (!P->launched) return fail(P->ctx, MCRT_ERR_NOT_READY, "plan was not launched");
  HIP_TRY(P->ctx, hipSetDevice(P->ctx->device));
  HIP_TRY(P->ctx, hipStreamSynchronize(P->ctx->stream));
  P->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - P->t_launch).count();
  return MCRT_OK;
}

extern "C" int mcrt_plan_kernel_ms(mcrt_plan* P, double* ms) {
  if (!P || !ms) return fail(nullptr, MCRT_ERR_INVALID, "plan_kernel_ms: NULL");
  if (!P->launched) return fail(P->ctx, MCRT_ERR_NOT_READY, "plan was not launched");
  float f = 0.f;
  HIP_TRY(P->ctx, hipEventElapsedTime(&f, P->ev0, P->ev1));
  *ms = f;
  return MCRT_OK;
}

extern "C" int mcrt_plan_read(mcrt_plan* P, double A[5], double* Irt, double* Icb, mcrt_stats* st) {
  if (!P) return fail(nullptr, MCRT_ERR_INVALID, "plan is NULL");
  mcrt_ctx* ctx = P->ctx;
  int rc = mcrt_plan_wait(P);
  if (rc) return rc;
  std::vector<double> hA(5), hrt(P->nrt), hcb(P->ncb);
  HIP_TRY(ctx, hipMemcpy(hA.data(), P->dA, sizeof(double) * 5, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(hrt.data(), P->dIrt, sizeof(double) * P->nrt, hipMemcpyDeviceToHost));
  if (P->ncb) HIP_TRY(ctx, hipMemcpy(hcb.data(), P->dIcb, sizeof(double) * P->ncb, hipMemcpyDeviceToHost));
  // simulate returns A/N, Irt/N, Icb/N (MonteCarloRadiativeTransfer.jl:78-84)
  const double N = (double)P->N;
  if (A) for (int i = 0; i < 5; ++i) A[i] = hA[i] / N;
  if (Irt) for (int64_t i = 0; i < P->nrt; ++i) Irt[i] = hrt[i] / N;
  if (Icb) for (int64_t i = 0; i < P->ncb; ++i) Icb[i] = hcb[i] / N;
  if (st) {
    unsigned long long c[mcrt::kCntN];
    HIP_TRY(ctx, hipMemcpy(c, P->dcnt, sizeof c, hipMemcpyDeviceToHost));
    st->rays = c[mcrt::kCntRays];
    st->events = c[mcrt::kCntEvents];
    st->rt_evals = c[mcrt::kCntRt];
    st->cb_evals = c[mcrt::kCntCb];
    st->cb_evals_single = c[mcrt::kCntCb1];
    st->rejections = c[mcrt::kCntRej];
    st->nonfinite = c[mcrt::kCntNonfinite];
    double ms = 0;
    mcrt_plan_kernel_ms(P, &ms);
    st->kernel_ms = ms;
    st->total_ms = P->total_ms;
    st->batches = P->precision == MCRT_PRECISION_FP64_PLAIN ? 0 : (uint64_t)P->coop.nslab;
    st->chunks = 0;
  }
  return MCRT_OK;
}

extern "C" int mcrt_plan_read_sigma(mcrt_plan* P, double A_sigma[5], double* Irt_sigma,
                                    double* Icb_sigma, int32_t* batches) {
  if (!P) return fail(nullptr, MCRT_ERR_INVALID, "plan is NULL");
  mcrt_ctx* ctx = P->ctx;
  if (P->precision == MCRT_PRECISION_FP64_PLAIN)
    return fail(ctx, MCRT_ERR_UNSUPPORTED,
                "standard errors come from the workgroup slabs of the cooperative forms "
                "(precision fp64 or mixed)");
  int rc = mcrt_plan_wait(P);
  if (rc) return rc;
  const int64_t nout = P->coop.nout;
  if (!P->dsig) {
    hipError_t e;
    P->dsig = static_cast<double*>(P->alloc(sizeof(double) * (size_t)nout, &e));
    if (!P->dsig) return fail(ctx, MCRT_ERR_NOMEM, "sigma buffer: %s", hipGetErrorString(e));
  }
  HIP_TRY(ctx, mcrt::coop_sigma(P->coop, ctx->stream, P->N, P->dsig));
  std::vector<double> h((size_t)nout);
  HIP_TRY(ctx, hipMemcpyAsync(h.data(), P->dsig, sizeof(double) * (size_t)nout,
                              hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (A_sigma) for (int i = 0; i < 5; ++i) A_sigma[i] = h[i];
  if (Irt_sigma) for (int64_t i = 0; i < P->nrt; ++i) Irt_sigma[i] = h[5 + i];
  if (Icb_sigma) for (int64_t i = 0; i < P->ncb; ++i) Icb_sigma[i] = h[5 + P->nrt + i];
  if (batches) *batches = P->coop.nslab;
  return MCRT_OK;
}

extern "C" int mcrt_plan_eval(mcrt_plan* P, int32_t what, int64_t n, const double* in, double* out) {
  if (!P) return fail(nullptr, MCRT_ERR_INVALID, "plan is NULL");
  mcrt_ctx* ctx = P->ctx;
  if (what < MCRT_EVAL_UPDATE_I || what > MCRT_EVAL_CB_PHASE)
    return fail(ctx, MCRT_ERR_INVALID, "plan_eval: unknown function %d", what);
  if (n < 0 || (n > 0 && (!in || !out))) return fail(ctx, MCRT_ERR_INVALID, "plan_eval: bad buffers");
  if (P->precision == MCRT_PRECISION_FP64_PLAIN &&
      (what == MCRT_EVAL_BIN_MATH || what == MCRT_EVAL_CB_PHASE))
    return fail(ctx, MCRT_ERR_UNSUPPORTED,
                "plan_eval: BIN_MATH and CB_PHASE are the cooperative forms' bin-loop functions");
  if (n == 0) return MCRT_OK;
  for (int64_t i = 0; what == MCRT_EVAL_DISTANCE && i < n; ++i) {
    const double idx = in[16 * i + 6];
    const int nl = P->geo_kind == MCRT_GEOMETRY_PLANE ? (P->precision == MCRT_PRECISION_FP64_PLAIN
                                                             ? P->plain.geo.nl : mcrt::kMaxLayers)
                                                      : 1;
    if (!(idx >= 0 && idx < nl && idx == (double)(int)idx))
      return fail(ctx, MCRT_ERR_INVALID, "plan_eval: layer index %g out of range", idx);
  }
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (P->precision != MCRT_PRECISION_FP64_PLAIN) {
    int rc = mcrt::coop_eval(P->coop, ctx->stream, what, n, in, out, &ctx->err);
    if (rc) return fail(ctx, rc, "plan_eval: %s", ctx->err.c_str());
    return MCRT_OK;
  }
  DevBuf din, dout;
  HIP_TRY(ctx, din.alloc(sizeof(double) * 16 * (size_t)n));
  HIP_TRY(ctx, dout.alloc(sizeof(double) * 4 * (size_t)n));
  HIP_TRY(ctx, hipMemcpyAsync(din.p, in, sizeof(double) * 16 * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, mcrt::launch_plain_eval(P->dplain, what, n, din.as<double>(), dout.as<double>(), ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(out, dout.p, sizeof(double) * 4 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return MCRT_OK;
}

extern "C" void mcrt_plan_destroy(mcrt_plan* P) {
  if (!P) return;
  (void)hipSetDevice(P->ctx->device);
  (void)hipStreamSynchronize(P->ctx->stream);
  delete P;
}

extern "C" int mcrt_simulate(mcrt_ctx* ctx, const mcrt_config* cfg, const mcrt_geometry* geom,
                             const double I0[4], uint64_t seed, int32_t precision, double A[5],
                             double* Irt, double* Icb, mcrt_stats* stats) {
  auto t0 = std::chrono::steady_clock::now();
  mcrt_plan* P = nullptr;
  int rc = mcrt_plan_create(ctx, cfg, geom, I0, seed, precision, &P);
  if (rc) return rc;
  rc = mcrt_plan_launch(P);
  if (!rc) rc = mcrt_plan_read(P, A, Irt, Icb, stats);
  mcrt_plan_destroy(P);
  if (!rc && stats)
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}
