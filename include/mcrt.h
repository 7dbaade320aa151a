/*
 * mcrt.h — C ABI of libmcrt_hip.so, the MI355X (gfx950) per-ray Monte Carlo
 * polarized radiative-transfer (RT) and coherent-backscattering (CB) transport.
 *
 * The library replaces the GPU branch of the reference's `simulate`
 * (src/MonteCarloRadiativeTransfer.jl:36-56, which launches `gpu_kernel`,
 * src/gpu.jl:177-296) and the scatterer table builders it depends on
 * (src/scatterers/mie.jl:36-44,129-216; src/utils.jl:263-287;
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

#define MCRT_ABI_VERSION 2

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

enum { MCRT_SCATTERER_RAYLEIGH = 0, MCRT_SCATTERER_MIE = 1, MCRT_SCATTERER_SPLINE = 2 };
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

/* A scatterer: RayleighScatterer (src/scatterers/rayleigh.jl:3-6),
 * MieScatterer (src/scatterers/mie.jl:3-15) or SplineScatterer
 * (src/scatterers/spline.jl:3-15).  Tables are read for Mie and Spline only,
 * F22 and F44 for Spline only (Mie's phase matrix has P22 = P11, P44 = P33).
 *
 * SplineScatterer: tables of one length n_theta on θ = range(0, π, n_theta),
 * ω and kl as given, csrn = calculate_csrn of F11 (spline.jl:19-24;
 * mcrt_calculate_csrn).  The reference's lookups (spline.jl:26-76) scale θ by
 * `self.NSPL`, a name that is not defined there, so every call throws
 * UndefVarError.  Here NSPL = n_theta - 1, the reading under which the lookup
 * spans the table; everything else is kept as written, including the phase
 * matrix's interpolation weights (msn+1-sn, sn-msn) on nodes msn, msn+1
 * (spline.jl:35-50), which extrapolate one node step as Mie's amplitude
 * lookup does (2·F[n_theta] at θ = π). */
typedef struct mcrt_scatterer {
  int32_t kind;           /* MCRT_SCATTERER_* */
  int32_t n_theta;        /* Mie, Spline: Nθ, length of F11..F44, S11, S22 */
  int32_t random_points;  /* Mie, Spline: csrn holds random_points+1 values */
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
  const double* F22;      /* Spline: [n_theta] */
  const double* F44;      /* Spline: [n_theta] */
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
  uint64_t batches;      /* batch means behind the standard errors: workgroup
                            slabs of a plan, batch slabs of a batched or served
                            solve (0: the plain form, which has none) */
  uint64_t chunks;       /* work chunks of a batched or served solve (0: plan) */
  double queue_ms;       /* served: submit to admission (0 elsewhere) */
  double wait_ms;        /* served: admission to the device start of the first
                            launch that ran one of its chunks (0 elsewhere);
                            kernel_ms is then that start to completion */
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
 * (src/MonteCarloRadiativeTransfer.jl:36-56).  The plan, batch and table
 * calls of one context are made from one host thread at a time (they share
 * the context's stream and its last-error message); a server's calls may come
 * from any thread (their messages are per calling thread). */
int mcrt_context_create(int device, mcrt_ctx** out);
/* Destroy plans and batches of a context before the context. */
void mcrt_context_destroy(mcrt_ctx* ctx);

/* Device-resident scatterer tables, shared by content: a plan or batch on the
 * context whose Mie scatterer has the same Nθ, random_points and tables as
 * one the context already holds uses those device tables instead of uploading
 * its own.  mcrt_table_upload uploads (or finds) the FP32 and FP64 tables of
 * a scatterer and pins them under *handle until mcrt_table_release; unpinned
 * tables stay cached, least recently used dropped first beyond 1 GiB.  A
 * Rayleigh scatterer has no tables (a valid handle that pins nothing).
 * Replaces run_rt's GPU preparation, which moves the scatterer tables to the
 * device (adapt to CuArray) once before its simulate calls
 * (src/MonteCarloRadiativeTransfer.jl:97-104 sphere, :194-199 plane). */
int mcrt_table_upload(mcrt_ctx* ctx, const mcrt_scatterer* sca, uint64_t* handle);
int mcrt_table_release(mcrt_ctx* ctx, uint64_t handle);
/* Device bytes held by the context's table cache and its lookup hits and
 * misses (plans, batches and uploads).  No reference counterpart. */
int mcrt_table_cache_stats(mcrt_ctx* ctx, int64_t* bytes, uint64_t* hits, uint64_t* misses);

/* ---- host table builders (FP64, no device use) ---------------------------- */

/* bhmie(x, m; Nθ, nextra, custom_nstop) — src/scatterers/mie.jl:129-216.
 * q[5] = (q_ext, q_sca, q_abs, q_back, asymm).  S1, S2: [2*n_theta]
 * interleaved complex.  F11..F34: [n_theta].  Any output may be NULL. */
int mcrt_bhmie(double x, double m_re, double m_im, int32_t n_theta, int32_t nextra,
               int32_t custom_nstop, double* q, double* S1, double* S2, double* F11,
               double* F12, double* F33, double* F34);

/* calculate_csrn(linear_interpolation(range(0, π, Nθ), F11), random_points)
 * — src/utils.jl:263-287 with src/scatterers/mie.jl:41-42.  csrn has
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

/* Launch shape of a cooperative plan (what the autotuner chooses): ray slots
 * per 512-thread workgroup (rays in flight, a multiple of 64; at most 256
 * mixed, 128 fp64; the event records in LDS are sized by it), workgroups
 * launched per CU (1..4), and where the CB amplitude table S is read (mixed
 * form with CB and a tabulated scatterer: > 0 or 0 staged in LDS when it fits
 * beside the tables and records, < 0 read through the caches from HBM; the
 * fp64 form always reads it from HBM).  0 selects the defaults: as many slots
 * as LDS allows (fewer when N is too small to give every CU a workgroup; S is
 * given up before slots when tables are large) and as many workgroups per CU
 * as LDS allows, at most 2.  The plain form ignores it.  No reference
 * counterpart. */
typedef struct mcrt_launch_shape {
  int32_t slots;
  int32_t workgroups_per_cu;
  int32_t stage_s;
  int32_t reserved;
} mcrt_launch_shape;

int mcrt_plan_create_shaped(mcrt_ctx* ctx, const mcrt_config* cfg, const mcrt_geometry* geom,
                            const double I0[4], uint64_t seed, int32_t precision,
                            uint64_t ray_begin, uint32_t instance, const mcrt_launch_shape* shape,
                            mcrt_plan** out);
/* The shape a plan was built with (workgroups: the launch's grid; stage_s 1
 * when S is staged in LDS, -1 when it is read from HBM). */
int mcrt_plan_shape(mcrt_plan* plan, mcrt_launch_shape* shape, int32_t* workgroups);
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

/* Known-answer evaluation of the plan's own device building blocks on n fixed
 * inputs: the functions the plan's transport kernel calls, in its arithmetic
 * (FP32 for MIXED, FP64 otherwise), with its scatterer tables and geometry.
 * in: n records of 16 doubles; out: n records of 4 doubles (unused words 0).
 * Blocking.  No reference counterpart: it exposes, for checks against the
 * reference's FP64 functions, what `simulate` only uses internally.
 *   UPDATE_I    in (cosθ, sinθ, cosϕ, sinϕ, I[4], k[3]) → update_I(I, k)
 *               (src/utils.jl:91-112) as the peel-off evaluates it
 *   UPDATE_E    in (frame[4], E₁ re, E₁ im, E₂ re, E₂ im, k[3]) → update_E
 *               (utils.jl:170-192) as the CB bins evaluate it
 *   UPDATE_E_PATH  the same, as the FP64 CB path state evaluates it
 *   DISTANCE    in (x[3], k[3], layer index) → (distance_to_boundary in the
 *               form's arithmetic, the same in FP64) (plane.jl:97-105,
 *               sphere.jl:69-73)
 *   PHASE       in (cosθ) → (P11, P12, P33, P34) of the peel-off's lookup
 *               (rayleigh.jl:12-17, mie.jl:53-71)
 *   AMPLITUDE   in (cosθ) → (S11 re, im, S22 re, im) (rayleigh.jl:20-23,
 *               mie.jl:76-92)
 *   SAMPLE      in (u) → cosθ drawn by inversion (rayleigh.jl:9-11, mie.jl:47-51)
 *   PHASE_TRANSPORT  in (cosθ) → the phase matrix as the transport's scatter
 *               evaluates it (utils.jl:298)
 *   KEPLER      in (M, e) → kepler_solver(M, e) (utils.jl:243-261)
 *   PHASE_DIAG  in (cosθ) → (P22, P44) of the peel-off's lookup, then (P22,
 *               P44) as the transport evaluates them (equal to P11, P33 except
 *               for a SplineScatterer, spline.jl:52-55)
 *   BIN_MATH    in (t, x, y, cosθ) → (e^{-t}, √x, 1/y, acos(cosθ)) as the
 *               peel-off and CB loops evaluate them (cpu.jl:20, sphere.jl:72,
 *               utils.jl:92-97, mie.jl:53; the arccos of a Mie or Spline plan's
 *               table lookup; y a normal number, x ≥ 0)
 *   CB_PHASE    in (r) → (cos 2πr, sin 2πr) as the CB bins evaluate the
 *               interference phase ΔΦ = 2πr (cpu.jl:94-99)
 *   (BIN_MATH and CB_PHASE: the cooperative forms only; -2 for FP64_PLAIN)
 * The table lookups use the scatterer of layer 1 (the plane's upper range and
 * CB, or the sphere's). */
enum {
  MCRT_EVAL_UPDATE_I = 0,
  MCRT_EVAL_UPDATE_E = 1,
  MCRT_EVAL_UPDATE_E_PATH = 2,
  MCRT_EVAL_DISTANCE = 3,
  MCRT_EVAL_PHASE = 4,
  MCRT_EVAL_AMPLITUDE = 5,
  MCRT_EVAL_SAMPLE = 6,
  MCRT_EVAL_PHASE_TRANSPORT = 7,
  MCRT_EVAL_KEPLER = 8,
  MCRT_EVAL_PHASE_DIAG = 9,
  MCRT_EVAL_BIN_MATH = 10,
  MCRT_EVAL_CB_PHASE = 11
};
int mcrt_plan_eval(mcrt_plan* plan, int32_t what, int64_t n, const double* in, double* out);

/* simulate with no device allocation: no reference counterpart; it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * outputs and normalisation as mcrt_simulate, for MCRT_PRECISION_FP64 and
 * MCRT_PRECISION_MIXED.  All device state of the solve (parameters, bin
 * constants, tables, batch slabs, ray-slot state) lives in a fixed 96 MiB
 * device array held by each device the library runs on until the process
 * ends; a solve that needs more is refused with MCRT_ERR_NOMEM and the bytes
 * it needs in mcrt_last_error.  The solve runs as a one-request batch
 * (mcrt_batch_*) and equals it up to the order of FP64 additions.  Blocking.
 * Calls that overlap on one device share the array and race: one call per
 * device at a time.  stats->kernel_ms is the host-timed launch sequence. */
int mcrt_simulate_static(mcrt_ctx* ctx, const mcrt_config* cfg, const mcrt_geometry* geom,
                         const double I0[4], uint64_t seed, int32_t precision, double A[5],
                         double* Irt, double* Icb, mcrt_stats* stats);

/* ---- batched solves ------------------------------------------------------- */

/* One solve of a batch: the arguments of simulate plus its stream position. */
typedef struct mcrt_request {
  const mcrt_config* cfg;
  const mcrt_geometry* geom;
  double I0[4];
  uint64_t seed;
  uint64_t ray_begin;   /* rays ray_begin .. ray_begin+N-1 of the stream */
  uint32_t instance;
  int32_t precision;    /* MCRT_PRECISION_FP64 or MCRT_PRECISION_MIXED */
  int64_t chunk_rays;   /* rays per work chunk; 0: the default, at least two
                           chunks when N >= 2.  Batches: ceil(N/64) rays within
                           512..8192 (throughput).  The server: about 4e7 bin
                           evaluations per chunk from the solve's bins and a
                           model of its events per ray, 64..8192 rays (a solve
                           is done no sooner than its longest chunk: latency). */
  int32_t slots;        /* ray slots per workgroup while a chunk of this solve
                           runs, and where its S table is read: slots and
                           stage_s of mcrt_launch_shape (what the autotuner
                           chooses; 0: the defaults) */
  int32_t stage_s;
} mcrt_request;

typedef struct mcrt_batch mcrt_batch;

/* Many solves (any mix of plane/sphere, Rayleigh/Mie, CB on/off, sizes) in
 * one persistent launch per precision: every solve is cut into chunks of
 * chunk_rays rays, the chunks of all solves form one queue ordered by their
 * estimated work, and each workgroup of the launch takes chunks until the
 * queue is empty.  A chunk's rays never share an iteration with another
 * chunk, so each solve gets the same result alone and in any batch, up to the
 * order of FP64 additions.  The sequence (clear, launch, reduce) is captured
 * in a HIP graph on the first mcrt_batch_launch and replayed afterwards.  The
 * request structs and what they point to are read during create only.  No
 * reference counterpart (the reference runs one simulate per call). */
int mcrt_batch_create(mcrt_ctx* ctx, int32_t n, const mcrt_request* reqs, mcrt_batch** out);
int mcrt_batch_launch(mcrt_batch* batch);  /* asynchronous on the context stream */
int mcrt_batch_wait(mcrt_batch* batch);
/* Results of solve i of the last launch, as mcrt_plan_read / _read_sigma. */
int mcrt_batch_read(mcrt_batch* batch, int32_t i, double A[5], double* Irt, double* Icb,
                    double A_sigma[5], double* Irt_sigma, double* Icb_sigma, mcrt_stats* stats);
/* Device time (ms) of the last launch: clear, transport and reduction. */
int mcrt_batch_kernel_ms(mcrt_batch* batch, double* ms);
/* Chunks in the queue and device bytes held by the batch (scatterer tables
 * shared through the context's table cache are not counted). */
int mcrt_batch_info(mcrt_batch* batch, int64_t* chunks, int64_t* device_bytes);
void mcrt_batch_destroy(mcrt_batch* batch);

/* ---- serving -------------------------------------------------------------- */

/* Ticket states. */
enum {
  MCRT_TICKET_QUEUED = 0,     /* waiting for device memory */
  MCRT_TICKET_RUNNING = 1,    /* admitted: its chunks are being launched */
  MCRT_TICKET_DONE = 2,
  MCRT_TICKET_CANCELLED = 3,
  MCRT_TICKET_FAILED = 4
};

typedef struct mcrt_server_stats {
  uint64_t submitted, completed, cancelled, failed;
  uint64_t launches;          /* persistent launches (windows) issued */
  uint64_t chunks;            /* tasks in those launches (chunks started or resumed) */
  uint64_t rays;              /* rays of the completed solves */
  uint64_t queued, running;   /* tickets now waiting / admitted */
  int64_t device_bytes_in_use, device_budget;
  int64_t table_cache_bytes;
  uint64_t table_hits, table_misses;
  double busy_ms;             /* wall time with a launch or a read-back in flight */
  uint64_t preemptions;       /* windows asked to stop early for waiting work */
  uint64_t resumed;           /* chunks resumed from a stopped window */
} mcrt_server_stats;

typedef struct mcrt_server mcrt_server;

/* A request loop on its own host thread and two HIP streams of the context's
 * device.  Submitted solves wait until their device memory (bins, batch
 * slabs, parameters) fits device_budget_bytes (<= 0: half of the free device
 * memory) and fewer than 1024 are admitted; Mie tables are kept resident
 * across requests in a cache keyed by their content (table_cache_bytes,
 * <= 0: 1 GiB).  Admitted solves of one precision share persistent launch
 * windows, one at a time: while other solves wait, the running window is
 * asked to stop after a time slice (at once when part of the device is idle),
 * its workgroups finish the iteration they are in and keep their chunks' rays
 * in flight, and the next window resumes those chunks round-robin with the
 * waiting ones.  Every solve's result is the same alone and in the mix, and
 * wherever the windows were cut, up to the order of FP64 additions.  No
 * reference counterpart. */
int mcrt_server_create(mcrt_ctx* ctx, int64_t device_budget_bytes, int64_t table_cache_bytes,
                       mcrt_server** out);
/* Queue a solve (the request and what it points to are copied). */
int mcrt_submit(mcrt_server* srv, const mcrt_request* req, uint64_t* ticket);
int mcrt_poll(mcrt_server* srv, uint64_t ticket, int32_t* state);
/* Block until the ticket is done, cancelled or failed (timeout_ms < 0: no
 * limit); -6 (not ready) on timeout. */
int mcrt_wait(mcrt_server* srv, uint64_t ticket, double timeout_ms, int32_t* state);
/* Results of a finished ticket (layouts and normalisation of mcrt_plan_read,
 * sigma as mcrt_plan_read_sigma; stats->total_ms = latency from submit), then
 * forget the ticket.  -5 for a cancelled ticket, the failure's status for a
 * failed one. */
int mcrt_result(mcrt_server* srv, uint64_t ticket, double A[5], double* Irt, double* Icb,
                double A_sigma[5], double* Irt_sigma, double* Icb_sigma, mcrt_stats* stats,
                double* latency_ms);
/* Stop a queued or running ticket (no effect once its last chunk finished). */
int mcrt_cancel(mcrt_server* srv, uint64_t ticket);
int mcrt_server_stats_get(mcrt_server* srv, mcrt_server_stats* out);
/* How long a window keeps the device while other work waits (default 8 ms;
 * 0: stop it as soon as work waits).  No reference counterpart. */
int mcrt_server_set_slice(mcrt_server* srv, double slice_ms);
/* Cancel what is left, wait for the device, free everything. */
void mcrt_server_destroy(mcrt_server* srv);

#ifdef __cplusplus
}
#endif

#endif /* MCRT_H */
