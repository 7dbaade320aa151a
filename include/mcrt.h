/*
 * mcrt.h — C ABI of libmcrt_hip.so, the MI355X (gfx950) per-ray Monte Carlo
 * polarized radiative-transfer (RT) and coherent-backscattering (CB) transport.
 *
 * The library replaces the GPU branch of the reference's `simulate`
 * (src/MonteCarloRadiativeTransfer.jl:36-56, which launches `gpu_kernel`,
 * src/gpu.jl:177-296) and the scatterer table builders it depends on
 * (src/scatterers/mie.jl:43-51,136-227; src/utils.jl:263-287;
 * FastGaussQuadrature.gausslegendre used at src/geometries/plane.jl:132 and
 * src/geometries/sphere.jl:86).  Every entry point is `extern "C"`, takes plain
 * pointers and sizes, and never throws: it returns an integer status, and
 * mcrt_last_error() gives the message.
 *
 * Layouts follow Julia's column-major arrays, 0-based here:
 *   Irt[s + 4*(j1 + Nθ*j2)]   s ∈ 0..3 (I,Q,U,V), j1 ∈ 0..Nθ-1, j2 ∈ 0..Nϕ-1
 *   Icb[s + 4*(j1 + Nθb*j2)]
 *   complex tables S11/S22: interleaved (re, im) doubles.
 * The caller owns every host buffer; the library copies inputs and keeps no
 * caller pointer after a call returns.
 */
#ifndef MCRT_H
#define MCRT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCRT_ABI_VERSION 1

/* Status codes.  MCRT_ERR_INVALID mirrors the reference's ArgumentError from
 * `validate` (src/geometries/plane.jl:52-62, src/geometries/sphere.jl:35-37). */
enum {
  MCRT_OK = 0,
  MCRT_ERR_INVALID = -1,
  MCRT_ERR_UNSUPPORTED = -2,
  MCRT_ERR_NOMEM = -3,
  MCRT_ERR_HIP = -4,
  MCRT_ERR_CANCELLED = -5,
  MCRT_ERR_NOT_READY = -6,
  MCRT_ERR_NO_DEVICE = -7
};

enum { MCRT_SCATTERER_RAYLEIGH = 0, MCRT_SCATTERER_MIE = 1 };
enum { MCRT_GEOMETRY_PLANE = 0, MCRT_GEOMETRY_SPHERE = 1 };

/* Arithmetic forms of the transport.
 *  FP64_PLAIN: the drop-in — one ray per thread, grid-stride over rays,
 *              FP64 everywhere, path history in HBM, O(n) CB re-walk,
 *              global FP64 atomics (the structure of src/gpu.jl:177-296).
 *  FP64:       persistent ray slots + workgroup-cooperative peel-off/CB bin
 *              evaluation, O(1) incremental CB path state, FP64 throughout.
 *  MIXED:      as FP64, with transport, sampling, rotations and bin
 *              evaluation in FP32; CB interference phase, CB path state and
 *              the bin accumulation in FP64. */
enum { MCRT_PRECISION_FP64_PLAIN = 0, MCRT_PRECISION_FP64 = 1, MCRT_PRECISION_MIXED = 2 };

/* A scatterer: RayleighScatterer (src/scatterers/rayleigh.jl:3-6) or
 * MieScatterer (src/scatterers/mie.jl:3-15).  Tables are read only for Mie. */
typedef struct mcrt_scatterer {
  int32_t kind;           /* MCRT_SCATTERER_* */
  int32_t n_theta;        /* Mie: Nθ, length of F11..F34, S11, S22 */
  int32_t random_points;  /* Mie: csrn holds random_points+1 values */
  int32_t reserved;
  double omega;           /* ω, single-scattering albedo */
  double kl;              /* kl, mean free path (also the CB phase scale) */
  const double* csrn;     /* [random_points+1] inverse CDF of cosθ */
  const double* F11;      /* [n_theta] on θ = range(0, π, Nθ) */
  const double* F12;
  const double* F33;
  const double* F34;
  const double* S11;      /* [2*n_theta] interleaved complex */
  const double* S22;
} mcrt_scatterer;

/* Layer (src/geometries/plane.jl:6-21).  d < 0 marks a semi-infinite layer
 * (Base.isinf(::Layer), plane.jl:23). */
typedef struct mcrt_layer {
  double d;
  double n;
  double tau;
  double omega;
  const mcrt_scatterer* scatterer;
} mcrt_layer;

/* PlaneGeometry (plane.jl:28-38) or SphereGeometry (sphere.jl:6-21). */
typedef struct mcrt_geometry {
  int32_t kind;                    /* MCRT_GEOMETRY_* */
  int32_t n_layers;                /* plane */
  double theta;                    /* plane incidence θ ∈ (π/2, π] */
  double phi;                      /* plane incidence ϕ ∈ [0, 2π) */
  const mcrt_layer* layers;        /* plane: [n_layers] */
  double R;                        /* sphere: actual radius */
  double n;                        /* sphere: host refractive index */
  double tau;                      /* sphere: optical radius */
  double omega;                    /* sphere: albedo */
  const mcrt_scatterer* scatterer; /* sphere */
} mcrt_geometry;

/* Config (src/config.jl:55-124).  The angle vectors are quadrature nodes and
 * weights (config.jl:93-99). */
typedef struct mcrt_config {
  int64_t number_of_rays;
  double minimum_intensity;
  int64_t maximum_scattering_times;
  int64_t maximum_threads;
  int32_t compute_cb;
  int32_t reserved;
  double tau0;                   /* τ₀ */
  int64_t n_phi;                 /* Nϕ */
  const double* phi;
  const double* cos_phi;
  const double* sin_phi;
  int64_t n_theta;               /* Nθ */
  const double* cos_theta;
  const double* sin_theta;
  const double* inorm;
  int64_t n_theta_b;             /* Nθb */
  const double* theta_b;
  const double* cos_theta_b;
  const double* sin_theta_b;
  int64_t n_phi_b;               /* Nϕb */
  const double* phi_b;
  const double* cos_phi_b;
  const double* sin_phi_b;
} mcrt_config;

/* Exact integer work counters of one solve. */
typedef struct mcrt_stats {
  uint64_t rays;
  uint64_t events;       /* scattering events (loop trips of cpu.jl:186) */
  uint64_t rt_evals;     /* peel-off direction evaluations (cpu.jl:11-28) */
  uint64_t cb_evals;     /* CB bin evaluations (cpu.jl:83-128, 131-158) */
  uint64_t cb_evals_single; /* of which single-scattering (n = 1) bins, cpu.jl:131-158 */
  uint64_t rejections;   /* resampled scatter/free-path draws (cpu.jl:206-234) */
  uint64_t nonfinite;    /* rays whose weight became NaN/Inf */
  double kernel_ms;      /* device time of the transport launch(es) */
  double total_ms;       /* wall time of the call */
} mcrt_stats;

typedef struct mcrt_ctx mcrt_ctx;
typedef struct mcrt_plan mcrt_plan;

/* ---- library / context --------------------------------------------------- */

/* No reference counterpart (version probe for the ccall binding). */
const char* mcrt_version(void);

/* Message of the last failed call on this context (ctx may be NULL: the
 * calling thread's last message).  No reference counterpart: the reference
 * raises Julia exceptions, the binding turns a non-zero status into
 * ErrorException(mcrt_last_error(ctx)). */
const char* mcrt_last_error(const mcrt_ctx* ctx);

/* Number of visible HIP devices (0 when none). */
int mcrt_device_count(int* count);

/* Create a context on a device: one HIP stream, table cache, scratch.
 * Replaces CUDA.jl's implicit device context used by simulate
 * (src/MonteCarloRadiativeTransfer.jl:36-56). */
int mcrt_context_create(int device, mcrt_ctx** out);
void mcrt_context_destroy(mcrt_ctx* ctx);

/* ---- host table builders (FP64, no device use) ---------------------------- */

/* bhmie(x, m; Nθ, nextra, custom_nstop) — src/scatterers/mie.jl:136-227.
 * q[5] = (q_ext, q_sca, q_abs, q_back, asymm).  S1, S2: [2*n_theta]
 * interleaved complex.  F11..F34: [n_theta].  Any output may be NULL. */
int mcrt_bhmie(double x, double m_re, double m_im, int32_t n_theta, int32_t nextra,
               int32_t custom_nstop, double* q, double* S1, double* S2, double* F11,
               double* F12, double* F33, double* F34);

/* calculate_csrn(linear_interpolation(range(0, π, Nθ), F11), random_points)
 * — src/utils.jl:263-287 with src/scatterers/mie.jl:48-49.  csrn has
 * random_points+1 values. */
int mcrt_calculate_csrn(const double* F11, int32_t n_theta, int32_t random_points,
                        double* csrn);

/* FastGaussQuadrature.gausslegendre(n): ascending nodes x[n], weights w[n]. */
int mcrt_gauss_legendre(int32_t n, double* x, double* w);

/* initial_weights(g, Nθ, Nϕ) — plane.jl:131-142, sphere.jl:85-90. */
int mcrt_initial_weights(int32_t geometry_kind, int32_t n_theta, int32_t n_phi,
                         double* cos_theta, double* sin_theta, double* inorm);

/* ---- transport ------------------------------------------------------------ */

/* simulate(cfg, g, I₀; use_gpu=true) — src/MonteCarloRadiativeTransfer.jl:26-85.
 * Blocking.  Outputs are normalised by number_of_rays exactly as the
 * reference's return value: A[5] = (Adt, Aabs, Aref, Atrans, Arem)/N,
 * Irt[4*Nθ*Nϕ]/N, Icb[4*Nθb*Nϕb]/N (fully overwritten).  Rays are numbered
 * 0..N-1; ray r draws its uniforms from Philox4x32-10 with key = seed and
 * counter = (r, draw index, instance 0).  stats may be NULL. */
int mcrt_simulate(mcrt_ctx* ctx, const mcrt_config* cfg, const mcrt_geometry* geom,
                  const double I0[4], uint64_t seed, int32_t precision, double A[5],
                  double* Irt, double* Icb, mcrt_stats* stats);

/* The same solve split into resident set-up, launch and read-back, so that a
 * caller can time the device work with all inputs already in HBM, or keep
 * several solves in flight.  No separate reference counterpart: together they
 * are simulate's GPU branch (MonteCarloRadiativeTransfer.jl:38-56). */
int mcrt_plan_create(mcrt_ctx* ctx, const mcrt_config* cfg, const mcrt_geometry* geom,
                     const double I0[4], uint64_t seed, int32_t precision,
                     mcrt_plan** out);
/* A plan over rays ray_begin .. ray_begin+N-1 of the stream (seed, instance),
 * N = cfg->number_of_rays; mcrt_plan_create is ray_begin = 0, instance = 0.
 * Plans over disjoint ranges of one stream are shards of one larger solve:
 * their N-weighted mean is that solve's output (up to summation order).  No
 * reference counterpart (the reference's RNG has no stream position). */
int mcrt_plan_create_range(mcrt_ctx* ctx, const mcrt_config* cfg, const mcrt_geometry* geom,
                           const double I0[4], uint64_t seed, int32_t precision,
                           uint64_t ray_begin, uint32_t instance, mcrt_plan** out);
int mcrt_plan_launch(mcrt_plan* plan);   /* asynchronous on the context stream */
int mcrt_plan_wait(mcrt_plan* plan);     /* blocks until the launch finished */
int mcrt_plan_read(mcrt_plan* plan, double A[5], double* Irt, double* Icb,
                   mcrt_stats* stats);
/* Device time (ms) of the transport kernel alone on the last launch, from HIP
 * events recorded on the context stream around it. */
int mcrt_plan_kernel_ms(mcrt_plan* plan, double* ms);
void mcrt_plan_destroy(mcrt_plan* plan);

/* Monte Carlo standard error of every output of the last launch (same
 * layouts and normalisation as mcrt_plan_read), by batch means over the
 * workgroup ray ranges: *batches ray ranges of near-equal size, each summed
 * separately on the device.  The reference reports no error estimate; this
 * is the optional `sigma` output of simulate.  Cooperative forms only
 * (MCRT_PRECISION_FP64, MCRT_PRECISION_MIXED; -2 for FP64_PLAIN).  NaN when
 * the solve ran as a single batch.  Any output may be NULL. */
int mcrt_plan_read_sigma(mcrt_plan* plan, double A_sigma[5], double* Irt_sigma,
                         double* Icb_sigma, int32_t* batches);

#ifdef __cplusplus
}
#endif

#endif /* MCRT_H */
