// mcrt_coop.hip — persistent, workgroup-cooperative transport (FP64 and mixed).
//
// The per-ray algorithm is cpu_kernel (src/cpu.jl:162-273) with collect_rt
// (cpu.jl:6-31) and collect_cb (cpu.jl:38-160); the work is laid out for
// CDNA4 instead of one-ray-per-thread:
//
//  * A 512-thread workgroup (8 wave64, two per SIMD) owns SLOTS ray slots
//    (256 in FP32, 128 in FP64).  One iteration advances every live slot by one
//    scattering event (scatter + rejection + absorption), writes a compact
//    event record to LDS (wave ballot + prefix over waves), refills terminated
//    slots from the workgroup's ray range, and repeats until the range is
//    exhausted.
//  * The peel-off (8640 directions per event at 180×48 bins) and the CB bins
//    are evaluated by all 512 threads: each thread owns a fixed set of bins
//    (register-resident direction constants), loops over the LDS event records
//    (broadcast reads) and accumulates its bins in registers, so no two threads
//    of a workgroup touch the same bin.  The escape sums Iref/Itrans that feed
//    back into each ray's weight are wave-reduced per (event, bin chunk) and
//    summed over waves in a fixed order.
//  * Bins are flushed after each iteration into the workgroup's private FP64
//    slab by atomic adds that return nothing (a slab takes at most ~10⁴
//    iteration partials, so its rounding stays near 1e-12 relative, far below
//    that of an FP32 partial); one reduction kernel sums the slabs in
//    workgroup order with Kahan compensation.  Results are bitwise
//    reproducible for a fixed launch.
//  * Mie F/csrn tables (and S in FP32) are staged in LDS, padded by one zero
//    node so that the interpolation needs no end-of-table branch.
//  * CB uses O(1) path state per ray in FP64: the forward field of
//    forward_E (utils.jl:199-216) advanced one step per event, and the
//    reversed path of backward_E (utils.jl:223-241) as a normalised running
//    2×2 complex product with a log scale, so no path history is stored.
//
// R is the transport / bin-evaluation type (double: FP64 form, float: mixed).
#include <algorithm>
#include <limits>
#include <map>
#include <mutex>
#include <cmath>
#include <cstring>
#include <vector>

#include "mcrt_coop.h"
#include "mcrt_internal.h"

// Translation units: build.py compiles this file once per part, so that the
// transport kernels build in parallel: 1 the host code and the reduction
// kernels, 2 the mixed form's plane kernels, 3 its sphere kernels, 4 its batch
// and known-answer kernels, 5 the FP64 form's kernels.  0 (no definition):
// everything in one unit.
#ifndef MCRT_COOP_PART
#define MCRT_COOP_PART 0
#endif
#define MCRT_PART(n) (MCRT_COOP_PART == 0 || MCRT_COOP_PART == (n))

namespace mcrt {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
constexpr int kRtRec = 20;  // R words of an RT event record
constexpr int kCbRec = 40;  // R words of a CB event record
constexpr int kCbDx = 4;    // doubles of a CB event record (x_a - x_b, τab)
constexpr int kBinW = 11;     // R words of a bin used by the evaluation
constexpr int kBinStride = 12;  // R words per bin in the bin tables
constexpr int64_t kMaxCoopBins = int64_t(1) << 26;  // bins per grid (RT or CB)
constexpr int kAcosTerms = 7;   // coefficients of the FP32 arccos polynomial (degree 6)

// Scatterer kind of a kernel instantiation: Rayleigh, Mie or Spline tables for
// every layer, or kAny: each table set's kind read at run time (a plane whose
// layers hold different scatterers, or any SplineScatterer), dispatched once
// per θ-range and per CB pass, outside the bin loops.
constexpr int kRay = 0, kMie = 1, kSpl = 2, kAny = 3;

// Per-precision launch shape: ray slots per workgroup and RT bins per
// register chunk.
template <typename R> struct Shape;
template <> struct Shape<float> {
  static constexpr int kSlots = 256;
  static constexpr int kChunk = 8;
  // peel-off chunk of the batch kernel's plane variants with tabulated
  // scatterers (Mie, general kind): at 8 bins their event loops spill 10-17
  // registers per event
  static constexpr int kChunkBatchPlane = 4;
  static constexpr bool kStageS = true;  // S table in LDS
  // ρ² below which a bin counts as aligned with the event direction (the
  // polar branch): FP32 rounding leaves ρ² ≈ 1e-13 for exactly (anti)parallel
  // directions, e.g. the θb = 0 CB bins at oblique incidence (ρ < 1e-5 rad)
  static constexpr float kPolar2 = 1e-10f;
};
template <> struct Shape<double> {
  static constexpr int kSlots = 128;
  static constexpr int kChunk = 2;
  static constexpr int kChunkBatchPlane = 2;
  static constexpr bool kStageS = false;
  static constexpr double kPolar2 = 1e-24;  // the reference's x² + y² < 1e-24
};

// Lookup constants of one scatterer table: θ-index = acos(μ)·(Nθ - 1)/π.
template <typename R> struct TabC {
  R acs[kAcosTerms];  // FP32 arccos coefficients times (Nθ - 1)/π, highest degree first
  R pis;              // π · (Nθ - 1)/π = Nθ - 1, the last node
  R scale;            // (Nθ - 1)/π
  int boff;           // Spline: offset of the (F22, F44) block, spl_b(Nθ)
};

template <typename R> struct CoopParams {
  int64_t N;
  int64_t maxscat;
  int slots;      // active ray slots per workgroup (multiple of 64, ≤ Shape<R>::kSlots)
  int stage_s;    // the S table is staged in LDS (FP32 with CB; else read from HBM)
  int rt_cut;     // 0: no peel-off distance can exceed τ₀ (sphere with 2τ < τ₀)
  int cb_cut;     // 0: no CB distance can exceed τ₀ (sphere with 4τ < τ₀)
  R min_int;
  R tau0;
  R tau0k;      // tau0 · ExpK<R>::k
  int compute_cb;
  int geo_kind;  // 0 plane, 1 sphere
  int nl;
  R tau[kMaxLayers];
  R omega[kMaxLayers];
  int isinf[kMaxLayers];
  double tau_d[kMaxLayers];
  double kl1;
  // Table sets: 0 is layer 1's scatterer (the sphere's; the plane's upper
  // θ-range, which only layer-1 events reach, and CB, which uses layer 1's
  // scatterer, cpu.jl:50), 1 the last layer's (the plane's lower θ-range,
  // which only last-layer events reach).  With one scatterer in every layer
  // set 1 is set 0 (lo_shared).  In R (transport, bins) and FP64 (CB path
  // state); F and S padded with one zero node.
  ScaView<R> sca;
  ScaView<R> sca_lo;
  ScaView<double> sca64;
  TabC<R> tc[2];
  int lo_shared;
  int sk;  // the kernel kind running the solve (kRay, kMie, kAny)
  // kAny: each layer's scatterer as the transport reads it (tables in HBM)
  ScaView<R> lsca[kMaxLayers];
  // RT bins: ranges of θ rows, [lower (transmission), upper (reflection)]
  int Nt, Np, Ntb, Npb;
  int nranges;
  int range_b[2], range_e[2], range_kind[2];  // kind 0: lower (A4), 1: upper (A3)
  const R* rt_bin;       // [Nt*Np][12]: k, Eh1 Eh2, Eh × k, inorm, 1/kz; bin = j1 + Nt*j2
  const R* cb_bin;       // [Ntb*Npb][12]: kf, Eh1 Eh2, Eh × kf, cos2ψ sin2ψ, 1/kfz
  const double* cb_kik;  // [Ntb*Npb][4]: k_i + k_f (FP64)
  R I0[4];
  R k0[3];
  double f0[4];   // raw incidence frame (cosθ₀, sinθ₀, cosϕ₀, sinϕ₀)
  double Ei[4];   // E_from_I(I₀): (E1 re, E1 im, E2 re, E2 im)
  double norm0;   // |E_i|
  double kf0[3];  // -k₀ (exact backscatter)
  uint64_t seed;
  uint64_t ray0;  // stream number of the first ray
  uint32_t instance;
  void* slot_mem;         // per workgroup: SLOTS Slot<R>, then SLOTS SlotAcc
  double* slabs;          // [blocks][2][nout]: Kahan sum, compensation
  int64_t nout;           // slab words: 5 + nrt + cb_groups·ncb (private slabs) or 5 + nrt + ncb
  int cb_groups;          // CB work items per bin: events e ≡ g (mod cb_groups)
  int64_t nrt, ncb;
  unsigned long long* counters;
};

namespace {

// Sum over the 64 lanes of a wave with DPP row operations (no LDS traffic):
// row_shr 1, 2, 4, 8 leave each row's total in its lane 15, row_bcast 15
// and 31 carry the totals into lane 63, which is broadcast by readlane.
template <int CTRL, int ROWS> MCRT_DEV int dpp_i(int x) {
  return __builtin_amdgcn_update_dpp(0, x, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS> MCRT_DEV float dpp_add(float v) {
  return v + __int_as_float(dpp_i<CTRL, ROWS>(__float_as_int(v)));
}
template <int CTRL, int ROWS> MCRT_DEV double dpp_add(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_i<CTRL, ROWS>((int)(b & 0xffffffffll));
  const int hi = dpp_i<CTRL, ROWS>((int)(b >> 32));
  return v + __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
MCRT_DEV float lane63(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }
MCRT_DEV double lane63(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <typename R> MCRT_DEV R wave_sum(R v) {
  v = dpp_add<0x111, 0xf>(v);  // row_shr:1
  v = dpp_add<0x112, 0xf>(v);  // row_shr:2
  v = dpp_add<0x114, 0xf>(v);  // row_shr:4
  v = dpp_add<0x118, 0xf>(v);  // row_shr:8
  v = dpp_add<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v = dpp_add<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return lane63(v);
}

// Kahan-compensated add of x into (s, c).
MCRT_DEV void kahan(double* s, double* c, double x) {
  double y = x - *c;
  double t = *s + y;
  *c = (t - *s) - y;
  *s = t;
}

// FP64 building blocks of the FP64 form's bin loops.  The library's are
// correctly rounded over their whole domains, with range and special-case
// handling (on gfx950: acos 93 VALU instructions, exp 42, sqrt 22, a division
// 12); these are branch-free, within a few ulp on the arguments the loops
// pass, and cost 17-25, 20, 10 and 5.
namespace f64 {
// 1/x for a normal, finite x ≠ 0: v_rcp_f64 and two Newton steps.
MCRT_DEV double rcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-x, y, 1.0);
  return __builtin_fma(y, e, y);
}
// √x for x ≥ 0: v_rsq_f64, one coupled Newton step on (√x, 1/(2√x)) and a
// final correction of √x by its residual; 0 below 2^-1000 (where the rsq of a
// flushed argument would give 0·∞).
MCRT_DEV double sqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double e = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, e, g);
  h = __builtin_fma(h, e, h);
  const double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return x > 0x1p-1000 ? g : 0.0;
}
// e^{-t}: -t = n·ln2 + r with n an integer and |r| ≤ ln2/2 (Cody and Waite's
// reduction, ln2 split at 32 bits so that n·ln2_hi is exact), e^r by its Taylor
// polynomial of degree 13 (remainder < 5e-18), times 2^n.  NaN stays NaN; the
// clamps keep 2^n within v_ldexp_f64's range (results below 2^-1075 are 0).
MCRT_DEV double exp_neg(double t) {
  double x = -t;
  x = x < -746.0 ? -746.0 : (x > 709.0 ? 709.0 : x);
  const double n = __builtin_rint(x * 1.4426950408889634);
  double r = __builtin_fma(n, -0.6931471803691238, x);  // ln2_hi = 0x1.62e42feep-1
  r = __builtin_fma(n, -1.9082149292705877e-10, r);    // ln2 - ln2_hi
  double p = 1.0 / 6227020800.0;                       // 1/13!
  p = __builtin_fma(p, r, 1.0 / 479001600.0);
  p = __builtin_fma(p, r, 1.0 / 39916800.0);
  p = __builtin_fma(p, r, 1.0 / 3628800.0);
  p = __builtin_fma(p, r, 1.0 / 362880.0);
  p = __builtin_fma(p, r, 1.0 / 40320.0);
  p = __builtin_fma(p, r, 1.0 / 5040.0);
  p = __builtin_fma(p, r, 1.0 / 720.0);
  p = __builtin_fma(p, r, 1.0 / 120.0);
  p = __builtin_fma(p, r, 1.0 / 24.0);
  p = __builtin_fma(p, r, 1.0 / 6.0);
  p = __builtin_fma(p, r, 0.5);
  p = __builtin_fma(p, r, 1.0);
  p = __builtin_fma(p, r, 1.0);
  return __builtin_amdgcn_ldexp(p, (int)n);
}
// acos(a) for a in [0, 1]: √(1-a)·p(a) with p the degree-17 weighted minimax
// polynomial of csrc/fit_acos.py (|acos error| ≤ 8.5e-17 rad before rounding),
// highest degree first.
constexpr int kAcosTerms64 = 18;
MCRT_DEV double acos01(double a) {
  constexpr double c[kAcosTerms64] = {
      -9.2665172311424016e-6, 9.0838956934216434e-5, -0.00042183485744450771,
      0.0012463190146607077, -0.0026692207269505534, 0.0045095056300445343,
      -0.0064400495955701362, 0.0082756854246340757, -0.010099913020109592,
      0.012182022890173891, -0.014879123878712378, 0.018671238874509465,
      -0.024374217369112181, 0.033681279022635642, -0.050792814208705652,
      0.089048622540930108, -0.21460183660249109, 1.5707963267948965};
  double p = c[0];
#pragma unroll
  for (int k = 1; k < kAcosTerms64; ++k) p = __builtin_fma(p, a, c[k]);
  return sqrt(1.0 - a) * p;
}
}  // namespace f64
// 1/x and √x of the bin loops (the FP32 hardware instructions; FP64: f64::).
MCRT_DEV float recip(float x) { return __builtin_amdgcn_rcpf(x); }
MCRT_DEV double recip(double x) { return f64::rcp(x); }
MCRT_DEV float bsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
MCRT_DEV double bsqrt(double x) { return f64::sqrt(x); }

// Table coordinate sn = acos(μ)·(Nθ-1)/π of a phase / amplitude lookup
// (mie.jl:53-55).  FP32: acos(a) = √(1-a)·p(a) on [0, 1] with p the degree-6
// weighted minimax polynomial of csrc/fit_acos.py (Remez exchange, |acos error|
// ≤ 8.8e-8 rad, below the FP32 rounding of θ and of μ itself), its coefficients
// pre-multiplied by (Nθ-1)/π, and acos(-a) = π - acos(a).  FP64: f64::acos01,
// the same form at degree 17.
// s2 ≥ 0 gives ρ² = 1 - μ² of the scattering plane: then 1 - a is formed as
// ρ²/(1 + a), which keeps θ accurate within 0.02° of 0 and π where μ itself
// rounds to ±1 in FP32 (the CB bins near exact backscattering, where the
// reference's lookups jump at θ = π); s2 < 0 (a constant at the call site)
// forms 1 - a from μ.
template <typename R> MCRT_DEV R theta_index(const TabC<R>& t, R mu, R s2 = R(-1));
template <> MCRT_DEV double theta_index(const TabC<double>& t, double mu, double) {
  const double r = f64::acos01(__builtin_fabs(clamp1(mu))) * t.scale;
  return mu < 0.0 ? t.pis - r : r;
}
template <> MCRT_DEV float theta_index(const TabC<float>& t, float mu, float s2) {
  const float a = __builtin_fabsf(mu);
  float c = t.acs[0];
#pragma unroll
  for (int k = 1; k < kAcosTerms; ++k) c = __builtin_fmaf(c, a, t.acs[k]);
  // 1 - a clamped to [0, 1] (|μ| can exceed 1 by rounding): the clamp bit of the subtraction
  const float om = s2 >= 0.0f ? s2 * __builtin_amdgcn_rcpf(1.0f + a)
                              : __builtin_amdgcn_fmed3f(1.0f - a, 0.0f, 1.0f);
  const float r = __builtin_amdgcn_sqrtf(om) * c;
  return mu < 0.0f ? t.pis - r : r;
}

MCRT_DEV float fmax_r(float a, float b) { return __builtin_fmaxf(a, b); }
MCRT_DEV double fmax_r(double a, double b) { return a > b ? a : b; }

// Split sn into the node index i0 ≤ imax and the fraction f = sn - i0.  In
// FP32, theta_index already lies in [0, imax] (its polynomial is positive, so
// r ≥ 0 and sn = r ≤ imax/2 or sn = imax - r), and at sn = imax the padded zero
// node takes the zero weight: no clamp is needed.
MCRT_DEV int node_of(float sn, float, float& f) {
  f = __builtin_amdgcn_fractf(sn);
  return (int)sn;
}
MCRT_DEV int node_of(double sn, double imax, double& f) {
  const double s = sn < imax ? sn : imax;
  const int i = (int)s;
  f = s - (double)i;
  return i;
}

// Phase matrix (P11, P12, P22, P33, P34, P44) of a bin lookup.
template <typename R> struct Ph { R p11, p12, p22, p33, p34, p44; };
// An aligned row of N table words, read as one vector (every table block starts
// on a 16-byte boundary and rows of 4 words sit at multiples of 4 words).
template <typename R, int N> struct alignas(N * sizeof(R)) Row { R v[N]; };

// Phase matrix at cosθ = mu: Rayleigh (rayleigh.jl:12-17), a lookup in the
// padded LDS table F linear in θ between nodes i0 and i0 + 1 (Mie,
// mie.jl:53-71, P22 = P11, P44 = P33), or the Spline lookup (spline.jl:26-58)
// with its weights (msn+1-sn, sn-msn) = (2-f, f-1) on nodes i0, i0 + 1 and
// six elements per node (the zero pad node gives 2·F at θ = π as written).
template <typename R, int SK>
MCRT_DEV Ph<R> phase_tab(const TabC<R>& t, const R* F, R mu, R s2 = R(-1)) {
  Ph<R> P;
  if constexpr (SK == kRay) {
    R m2 = mu * mu;
    P.p11 = R(0.75) * (R(1) + m2);
    P.p12 = R(-0.75) * (R(1) - m2);
    P.p33 = R(1.5) * mu;
    P.p34 = R(0);
    P.p22 = P.p11;
    P.p44 = P.p33;
  } else if constexpr (SK == kMie) {
    R f;
    const int i0 = node_of(theta_index(t, mu, s2), t.pis, f);
    const R* a = F + 4 * i0;
    P.p11 = a[0] + f * (a[4] - a[0]);
    P.p12 = a[1] + f * (a[5] - a[1]);
    P.p33 = a[2] + f * (a[6] - a[2]);
    P.p34 = a[3] + f * (a[7] - a[3]);
    P.p22 = P.p11;
    P.p44 = P.p33;
  } else {
    static_assert(SK == kSpl, "phase_tab: a table kind");
    R f;
    const int i0 = node_of(theta_index(t, mu, s2), t.pis, f);
    const R g = R(1) - f;
    // whole rows: (F11, F12, F33, F34) of nodes i0, i0 + 1 and their (F22, F44)
    // (left to itself the compiler pairs a[c] with a[c + 4] into 2 × 4-byte
    // reads, 6 LDS instructions per lookup instead of 3-4)
    const Row<R, 4> a0 = *(const Row<R, 4>*)(F + 4 * i0);
    const Row<R, 4> a1 = *(const Row<R, 4>*)(F + 4 * i0 + 4);
    const Row<R, 2> b0 = *(const Row<R, 2>*)(F + t.boff + 2 * i0);
    const Row<R, 2> b1 = *(const Row<R, 2>*)(F + t.boff + 2 * i0 + 2);
    P.p11 = a0.v[0] + g * (a0.v[0] - a1.v[0]);
    P.p12 = a0.v[1] + g * (a0.v[1] - a1.v[1]);
    P.p22 = b0.v[0] + g * (b0.v[0] - b1.v[0]);
    P.p33 = a0.v[2] + g * (a0.v[2] - a1.v[2]);
    P.p34 = a0.v[3] + g * (a0.v[3] - a1.v[3]);
    P.p44 = b0.v[1] + g * (b0.v[1] - b1.v[1]);
  }
  return P;
}

// Amplitude matrix diag(S11, S22) (rayleigh.jl:20-23; mie.jl:76-92 and
// spline.jl:60-76, the same lookup, with its weights (msn+1-sn, sn-msn) =
// (2-f, f-1) on nodes i0, i0 + 1: the zero pad node reproduces the skipped
// term at θ = π).
template <typename R, int SK>
MCRT_DEV void amp_tab(const TabC<R>& t, const R* S, R mu, Cx<R>& S11, Cx<R>& S22, R s2 = R(-1)) {
  if constexpr (SK == kRay) {
    S11 = {R(0), R(1)};
    S22 = {mu * R(0), mu};
  } else {
    R f;
    const int i0 = node_of(theta_index(t, mu, s2), t.pis, f);
    const R g = R(1) - f;
    const Row<R, 4> a0 = *(const Row<R, 4>*)(S + 4 * i0);  // nodes i0, i0 + 1 as rows
    const Row<R, 4> a1 = *(const Row<R, 4>*)(S + 4 * i0 + 4);
    S11 = {a0.v[0] + g * (a0.v[0] - a1.v[0]), a0.v[1] + g * (a0.v[1] - a1.v[1])};
    S22 = {a0.v[2] + g * (a0.v[2] - a1.v[2]), a0.v[3] + g * (a0.v[3] - a1.v[3])};
  }
}

// ------------------------------------------------------------- geometry (R)
template <typename R> struct Geo {
  const CoopParams<R>* p;
  MCRT_DEV R dist(V3<R> x, V3<R> k, int idx) const {  // distance_to_boundary
    if (p->geo_kind == 0) {
      if (k.z > R(0)) return -x.z / k.z;
      if (!p->isinf[idx]) return (x.z + p->tau[idx]) / -k.z;
      return R(INFINITY);
    }
    R kx = dot(k, x), xx = dot(x, x);
    return -kx + M<R>::sqrt(kx * kx + p->tau[0] * p->tau[0] - xx);
  }
};

// FP64 geometry for the CB path state.
template <typename R> MCRT_DEV double dist64(const CoopParams<R>& p, V3<double> x, V3<double> k, int idx) {
  if (p.geo_kind == 0) {
    if (k.z > 0) return -x.z / k.z;
    if (!p.isinf[idx]) return (x.z + p.tau_d[idx]) / -k.z;
    return INFINITY;
  }
  double kx = dot(k, x), xx = dot(x, x);
  return -kx + sqrt(kx * kx + p.tau_d[0] * p.tau_d[0] - xx);
}

// Frame rotation matrix of ray_to_norm (utils.jl:30-34): ray_to_norm(v) = M v,
// norm_to_ray(v) = Mᵀ v.  Stored as 8 words (M21 = 0); its columns are the
// frame's θ̂, ϕ̂ and k.
template <typename R> MCRT_DEV void frame_matrix(Fr<R> f, R m[8]) {
  m[0] = f.ct * f.cp; m[1] = -f.sp; m[2] = f.st * f.cp;
  m[3] = f.ct * f.sp; m[4] = f.cp;  m[5] = f.st * f.sp;
  m[6] = -f.st;       m[7] = f.ct;
}

// Scattering plane of a bin direction kb (basis Eh, G = Eh × kb, both
// precomputed per bin) seen from an event frame m, as needed by update_I
// (utils.jl:91-112) and update_E (utils.jl:170-192).  With r = norm_to_ray(kb)
// and ρ² = r₁² + r₂² = 1 - μ²:
//   cosϕₖ = r₁/ρ, sinϕₖ = r₂/ρ, μ = r₃,
//   cosψ = Eh·Ehₖ = k·(Eh × kb)/ρ = pn/ρ   (Binet-Cauchy, θ̂ × ϕ̂ = k),
//   Eh·Evₖ = -(k·Eh)/ρ = -w/ρ             (Eh ⟂ kb, |kb| = 1),
// so that every rotation of a peel-off is a ratio over the one factor
// inv = 1/ρ² and needs no square root.  The polar branch of
// spherical_from_cartesian (utils.jl:12-14, x² + y² < 1e-24: ϕₖ = 0,
// θₖ ∈ {0, π}; here ρ² < Shape<R>::kPolar2) is taken by a wave-uniform test,
// since exact alignment of a bin with an event direction is rare (a
// divergent branch that is skipped when no lane of the wave takes it).
// c2 = rx² - ry² = ρ²·cos2ϕₖ and r2 = ρ² are kept for the Stokes rotation, s2
// (ρ² as computed, -1 in the polar case) for the lookups of the CB bins.
template <typename R> struct Plane { R mu, rx, ry, pn, w, c2, r2, inv, s2; };

MCRT_DEV float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
MCRT_DEV double fma_r(double a, double b, double c) { return ::fma(a, b, c); }

// The general case; returns ρ² (q.inv is not set).
template <typename R>
MCRT_DEV R plane_raw(Plane<R>& q, const R* m, R kx, R ky, R kz, R e1, R e2, R g0, R g1, R g2) {
  q.rx = m[0] * kx + m[3] * ky + m[6] * kz;
  q.ry = m[1] * kx + m[4] * ky;
  q.mu = m[2] * kx + m[5] * ky + m[7] * kz;
  q.pn = m[2] * g0 + m[5] * g1 + m[7] * g2;
  q.w = m[2] * e1 + m[5] * e2;
  const R ry2 = q.ry * q.ry;
  q.c2 = fma_r(q.rx, q.rx, -ry2);
  q.r2 = fma_r(q.rx, q.rx, ry2);
  q.s2 = q.r2;
  return q.r2;
}

// The polar case (ρ² < Shape<R>::kPolar2): Ehₖ = -ϕ̂, Evₖ = sign(μ)·θ̂; returns ρ² := 1.
template <typename R>
MCRT_DEV R plane_polar(Plane<R>& q, const R* m, R e1, R e2) {
  const R u = e1 * m[0] + e2 * m[3], v = e1 * m[1] + e2 * m[4];
  q.rx = R(1);
  q.ry = R(0);
  q.pn = -v;
  q.w = q.mu > R(0) ? -u : (q.mu < R(0) ? u : R(0));
  q.c2 = R(1);
  q.r2 = R(1);
  q.s2 = R(-1);  // the lookup takes θ from μ = ±1, as the reference's polar case
  return R(1);
}

template <typename R>
MCRT_DEV Plane<R> scat_plane(const R* m, R kx, R ky, R kz, R e1, R e2, R g0, R g1, R g2) {
  Plane<R> q;
  R r2 = plane_raw(q, m, kx, ky, kz, e1, e2, g0, g1, g2);
  if (r2 < Shape<R>::kPolar2) r2 = plane_polar(q, m, e1, e2);
  q.inv = recip(r2);
  return q;
}

// Peel-off Stokes vector of one bin before attenuation: update_I
// (utils.jl:91-112) = R(ψ)·P(μ)·R(-ϕₖ)·I on the scattering plane q.  I is the
// event's Stokes vector, J = (2·I₂, -2·I₃).  Returns (o0, o1·ρ², o2·ρ², o3):
// the middle pair still carries the factor ρ² (multiply by q.inv).
template <typename R, int SK, bool ACC = false>
MCRT_DEV void peel_I(const TabC<R>& t, const R* F, const R* I, const R* J,
                     const Plane<R>& q, R o[4]) {
  const Ph<R> P = phase_tab<R, SK>(t, F, q.mu, ACC ? q.s2 : R(-1));
  const R c2 = q.c2;                      // ρ²·cos2ϕₖ
  const R sh = q.rx * q.ry;                // -ρ²·sin2ϕₖ / 2
  const R i1 = (c2 * I[1] + sh * J[1]) * q.inv;
  const R i2 = (sh * J[0] + c2 * I[2]) * q.inv;
  const R a0 = P.p11 * I[0] + P.p12 * i1;
  const R a1 = P.p12 * I[0] + P.p22 * i1;
  const R a2 = P.p33 * i2 + P.p34 * I[3];
  const R a3 = -P.p34 * i2 + P.p44 * I[3];
  // ρ²·cos2ψ = pn² - w² = ρ² - 2w² (pn² + w² = ρ²: cos²ψ + sin²ψ = 1), ρ²·sin2ψ = 2w·pn
  const R w2 = q.w + q.w;
  const R c2p = fma_r(-w2, q.w, q.r2);
  const R s2p = w2 * q.pn;
  o[0] = a0;
  o[1] = c2p * a1 + s2p * a2;
  o[2] = -s2p * a1 + c2p * a2;
  o[3] = a3;
}

// update_E (utils.jl:170-192) on the scattering plane q: E ← Rψ·S(μ)·Rϕₖ·E,
// returned times ρ²·sc (sc = q.inv gives update_E itself); the real factor sc
// (e.g. an attenuation) rides on the two coefficients of the ψ rotation.
// ACC: the lookup's θ from ρ² (theta_index).
template <typename R, int SK, bool ACC = false>
MCRT_DEV void peel_E(const TabC<R>& t, const R* S, const Plane<R>& q, R sc, Cx<R>& E1,
                     Cx<R>& E2) {
  Cx<R> S11, S22;
  amp_tab<R, SK>(t, S, q.mu, S11, S22, ACC ? q.s2 : R(-1));
  const Cx<R> f1 = S11 * (E1 * q.rx + E2 * q.ry);
  const Cx<R> f2 = S22 * (E2 * q.rx + E1 * -q.ry);
  const R pn = q.pn * sc, w = q.w * sc;
  E1 = f1 * pn + f2 * -w;
  E2 = f1 * w + f2 * pn;
}

// Attenuation scale: positions and path lengths of the RT records are stored
// times kExpK so that the FP32 form evaluates e^{-t} as one exp2.
template <typename R> struct ExpK;
template <> struct ExpK<float> {
  static constexpr float k = 1.4426950408889634f;  // log₂e
  static MCRT_DEV float exp_neg(float t) { return __builtin_amdgcn_exp2f(-t); }
};
template <> struct ExpK<double> {
  static constexpr double k = 1.0;
  static MCRT_DEV double exp_neg(double t) { return f64::exp_neg(t); }
};

// cos/sin of the large FP64 CB phase, given in revolutions (kl·(k_i + k_f)
// carries the factor 1/2π), so that its reduction to [-1/2, 1/2] revolutions,
// rev - rint(rev), is exact.  FP32 evaluates the reduced phase with
// v_sin_f32/v_cos_f32.  FP64 splits it further into quarter turns q and
// h ∈ [-1/2, 1/2] (4·f - rint(4·f), exact) and evaluates sin(πh/2) and
// cos(πh/2) by their Taylor polynomials (through h¹⁵ and h¹⁶: remainders
// below 7e-17), rotated by q quarter turns; the library sincos of the phase in
// radians costs 154 VALU instructions with a branch for large arguments.
constexpr double kRevPerRadian = 0.15915494309189535;  // 1/2π
MCRT_DEV void phase_cs(double rev, double& c, double& s) {
  const double g = 4.0 * (rev - __builtin_rint(rev));
  const double q = __builtin_rint(g);
  const double h = g - q, z = h * h;
  // (π/2)^(2k+1)/(2k+1)! and (π/2)^(2k)/(2k)! with alternating signs
  double ps = -6.688035109811468e-10;
  ps = __builtin_fma(ps, z, 5.692172921967927e-08);
  ps = __builtin_fma(ps, z, -3.598843235212085e-06);
  ps = __builtin_fma(ps, z, 0.00016044118478735983);
  ps = __builtin_fma(ps, z, -0.004681754135318688);
  ps = __builtin_fma(ps, z, 0.07969262624616705);
  ps = __builtin_fma(ps, z, -0.6459640975062463);
  ps = __builtin_fma(ps, z, 1.5707963267948966);
  const double sn = h * ps;
  double pc = 6.565963114979473e-11;
  pc = __builtin_fma(pc, z, -6.386603083791852e-09);
  pc = __builtin_fma(pc, z, 4.710874778818172e-07);
  pc = __builtin_fma(pc, z, -2.5202042373060607e-05);
  pc = __builtin_fma(pc, z, 0.0009192602748394266);
  pc = __builtin_fma(pc, z, -0.02086348076335296);
  pc = __builtin_fma(pc, z, 0.25366950790104803);
  pc = __builtin_fma(pc, z, -1.2337005501361697);
  pc = __builtin_fma(pc, z, 1.0);
  const int qi = (int)q & 3;  // q ∈ {-2..2}: quarter turns mod 4
  s = qi == 0 ? sn : (qi == 1 ? pc : (qi == 2 ? -sn : -pc));
  c = qi == 0 ? pc : (qi == 1 ? -sn : (qi == 2 ? -pc : sn));
}
MCRT_DEV void phase_cs(double rev, float& c, float& s) {
  float f = (float)(rev - rint(rev));      // in [-1/2, 1/2] revolutions
  s = __builtin_amdgcn_sinf(f);            // v_sin_f32 takes revolutions
  c = __builtin_amdgcn_cosf(f);
}

// The kind word of a CB record (0 none, 1 n = 1, 2 n > 1) as integer bits in
// an R word, read without a conversion.
template <typename R> MCRT_DEV R kind_word(int k);
template <> MCRT_DEV float kind_word<float>(int k) { return __int_as_float(k); }
template <> MCRT_DEV double kind_word<double>(int k) { return __longlong_as_double((long long)k); }
MCRT_DEV int kind_of(float w) { return __float_as_int(w); }
MCRT_DEV int kind_of(double w) { return (int)__double_as_longlong(w); }

MCRT_DEV void lmap_normalise(Lmap<double>& Q, double& lq) {
  double m = fmax(fmax(fmax(fabs(Q.a.re), fabs(Q.a.im)), fmax(fabs(Q.b.re), fabs(Q.b.im))),
                  fmax(fmax(fabs(Q.c.re), fabs(Q.c.im)), fmax(fabs(Q.d.re), fabs(Q.d.im))));
  if (m > 0 && isfinite(m)) {
    int e;
    frexp(m, &e);
    double sc = ldexp(1.0, -e);  // exact power-of-two rescale
    Q.a = Q.a * sc; Q.b = Q.b * sc; Q.c = Q.c * sc; Q.d = Q.d * sc;
    lq += e * 0.6931471805599453;
  }
}

// Per-slot ray state (HBM between phases): the transport part, copied into
// registers by the transport and update phases, and the CB path state, read
// and written in place by cb_event alone (a whole-record copy kept ~300 B per
// lane in scratch memory across the transport).
template <typename R> struct RaySt {
  int live;
  int idx, idxp;
  int64_t n;
  uint64_t ray;
  uint32_t draw;
  R I[4], Ip[4];
  V3<R> k, kp, x, xp;
};
struct CbSt {  // FP64
  int idx1;         // layer of the first scattering point
  V3<double> x1d;   // first scattering point
  V3<double> kprev; // previous incoming direction
  Fr<double> ksf;   // frame of the last forward step
  Cx<double> Ea1, Ea2;
  double enorma;
  Lmap<double> Q;
  double lq;
  Fr<double> kbs;   // frame of -k₂
};
template <typename R> struct Slot : RaySt<R>, CbSt {};

// Per-slot running sums of the energy budget and the work counters, kept in
// HBM beside the slot state (not in registers across the bin loops).
struct SlotAcc {
  double A[5];
  unsigned long long cnt[kCntN];
};

// The slot state and accumulators of workgroup `block`: SLOTS Slot<R>
// records followed by SLOTS SlotAcc records.
template <typename R> MCRT_HD size_t slot_block_bytes() {
  return (sizeof(Slot<R>) + sizeof(SlotAcc)) * (size_t)Shape<R>::kSlots;
}
template <typename R> MCRT_DEV Slot<R>* slot_block(void* base, int block) {
  return (Slot<R>*)((unsigned char*)base + slot_block_bytes<R>() * (size_t)block);
}

// One pass of the peel-off (cpu.jl:11-28) over CH register-resident bins per
// thread and all E events of the iteration.  Thread tid owns the bins
// lb = tid + kThreads·(c0 + i) of the θ-range [jb, jb + nrow) × Nϕ; the escape
// sum of each event is wave-reduced and added to esc in chunk order by the
// lane of the event within its block of 64.
// RT record: m[0..7], I[8..11], J[12..13], K·x[14..16], [17] K·τ of the layer
// (plane) or K²·(τ² - x·x) (sphere), [18] 1 or 0: near the upper (first) boundary,
// [19] 1 or 0: near the lower (last) boundary; K = ExpK<R>::k.
template <int CH, typename R, int GEO, int SK, bool CUT>
MCRT_DEV void rt_pass(const CoopParams<R>& p, const TabC<R>& t, const R* F, const R* rtrec, R* esc,
                      double* slab, int E, int tid, int lane, int wave, int jb, int nrow, int nb,
                      int c0, int fw, int slot2) {
  R bk[CH][kBinW];
  int bidx[CH];
  R acc[CH][4];
#pragma unroll
  for (int i = 0; i < CH; ++i) {
    const int lb = tid + kThreads * (c0 + i);
    const bool valid = lb < nb;
    const int j1 = jb + lb % nrow, j2 = lb / nrow;
    const int ob = j1 + p.Nt * j2;
    bidx[i] = valid ? ob : -1;
    const R* src = p.rt_bin + kBinStride * (valid ? ob : 0);
#pragma unroll
    for (int c = 0; c < kBinW; ++c) bk[i][c] = src[c];
    if (!valid) bk[i][8] = R(0);  // zero weight
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[i][c] = R(0);
  }
  const R tau0k = p.tau0k;
  // events in blocks of 64: lane l holds the near flag and, at the end, the
  // wave's escape sum of event e0 + l, added to esc by that lane
  for (int e0 = 0; e0 < E; e0 += 64) {
    const int el = e0 + lane;
    const unsigned long long near = __ballot(el < E && rtrec[kRtRec * el + fw] != R(0));
    R tot = R(0);
    const int e1 = min(E, e0 + 64);
    for (int e = e0; e < e1; ++e) {
      if (!((near >> (e - e0)) & 1ull)) continue;  // not near the range's boundary
      const R* rec = rtrec + kRtRec * e;
      // the scattering planes of all CH bins first, with one (rarely taken)
      // branch for the polar case, so that the lookups and Stokes algebra of the
      // CH bins form one straight-line block the scheduler can interleave
      Plane<R> qs[CH];
      R r2[CH];
      bool polar = false;
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        r2[i] = plane_raw(qs[i], rec, bk[i][0], bk[i][1], bk[i][2], bk[i][3], bk[i][4], bk[i][5],
                          bk[i][6], bk[i][7]);
        polar |= r2[i] < Shape<R>::kPolar2;
      }
      if (polar) {
#pragma unroll
        for (int i = 0; i < CH; ++i)
          if (r2[i] < Shape<R>::kPolar2) r2[i] = plane_polar(qs[i], rec, bk[i][3], bk[i][4]);
      }
      R part = R(0);
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        Plane<R>& q = qs[i];
        q.inv = recip(r2[i]);
        R o[4];
        peel_I<R, SK>(t, F, rec + 8, rec + 12, q, o);
        R t;  // K·distance_to_boundary
        if (GEO == 0) {  // plane.jl:97-105
          t = (bk[i][2] > R(0) ? rec[16] : rec[16] + rec[17]) * -bk[i][9];
        } else {         // sphere.jl:69-73
          const R kx = bk[i][0] * rec[14] + bk[i][1] * rec[15] + bk[i][2] * rec[16];
          t = bsqrt(kx * kx + rec[17]) - kx;  // rec[17] ≥ 0: the radicand is too
        }
        // the τ₀ cut-off (cpu.jl:20); CUT = false when no distance can exceed it
        const R w = (!CUT || t <= tau0k) ? bk[i][8] * ExpK<R>::exp_neg(t) : R(0);
        const R wi = w * q.inv;
        acc[i][0] += o[0] * w;
        acc[i][1] += o[1] * wi;
        acc[i][2] += o[2] * wi;
        acc[i][3] += o[3] * w;
        part += o[0] * w;
      }
      const R sum = wave_sum(part);
      if (lane == e - e0) tot = sum;
    }
    if (el < E) esc[(el * kWaves + wave) * 2 + slot2] += tot;
  }
  // flush into the slab by FP64 atomic adds that return nothing, so the
  // workgroup does not wait for a read of the slab: into a workgroup-private
  // slab (a bin is owned by one thread, so the adds to an address arrive in
  // program order and the sums are reproducible) or a slab shared by chunks
#pragma unroll
  for (int i = 0; i < CH; ++i) {
    if (bidx[i] < 0) continue;
    const int64_t o = 5 + 4 * (int64_t)bidx[i];
#pragma unroll
    for (int c = 0; c < 4; ++c) unsafeAtomicAdd(slab + o + c, (double)acc[i][c]);
  }
}

// One θ-range of the peel-off with the table set (t, F): full register chunks
// of CH bins, then the remainder in passes of 4, 2 and 1.
template <int CH, typename R, int GEO, int SK, bool CUT>
MCRT_DEV void rt_range(const CoopParams<R>& p, const TabC<R>& t, const R* F, const R* rtrec, R* esc,
                       double* slab, int E, int tid, int lane, int wave, int rg) {
  const int jb = p.range_b[rg], je = p.range_e[rg];
  const int nrow = je - jb;
  const int nb = nrow * p.Np;  // bins of this range
  const int per = (nb + kThreads - 1) / kThreads;
  const int fw = p.range_kind[rg] == 1 ? 18 : 19;  // record word of the near flag
  const int slot2 = p.range_kind[rg] == 1 ? 0 : 1;
  int c0 = 0;
  // with chunks of 8, a remainder of one bin joins the last chunk (a pass of
  // 9: e.g. 17 bins per thread at 180 × 48 run as 9 + 8, and 9 as one pass),
  // since every pass pays the per-event overhead (record reads, reduction).
  // Not for a plane's tabulated scatterers: their 9-bin pass spills ~30
  // registers per event around the polar block (8 + 1 is 8% faster there)
  if constexpr (CH == 8 && (GEO == 1 || SK == kRay)) {
    if (per % 8 == 1 && per >= 9) {
      rt_pass<9, R, GEO, SK, CUT>(p, t, F, rtrec, esc, slab, E, tid, lane, wave, jb, nrow, nb, c0,
                                  fw, slot2);
      c0 += 9;
    }
  }
  for (; c0 + CH <= per; c0 += CH)
    rt_pass<CH, R, GEO, SK, CUT>(p, t, F, rtrec, esc, slab, E, tid, lane, wave, jb, nrow, nb, c0,
                                 fw, slot2);
  if (CH > 4 && c0 + 4 <= per) {
    rt_pass<4, R, GEO, SK, CUT>(p, t, F, rtrec, esc, slab, E, tid, lane, wave, jb, nrow, nb, c0,
                                fw, slot2);
    c0 += 4;
  }
  if (c0 + 2 <= per) {
    rt_pass<2, R, GEO, SK, CUT>(p, t, F, rtrec, esc, slab, E, tid, lane, wave, jb, nrow, nb, c0,
                                fw, slot2);
    c0 += 2;
  }
  if (c0 < per)
    rt_pass<1, R, GEO, SK, CUT>(p, t, F, rtrec, esc, slab, E, tid, lane, wave, jb, nrow, nb, c0,
                                fw, slot2);
}

// All θ-ranges of the peel-off.  The upper range (reflection; the sphere's
// only range) uses table set 0 (F0), the plane's lower range table set 1 (F1).
// kAny dispatches on the set's kind once per range: a sphere's set is a
// Spline (Rayleigh and Mie spheres have kernels of their own), a plane's any.
template <int CH, typename R, int GEO, int SK, bool CUT>
MCRT_DEV void rt_ranges(const CoopParams<R>& p, const R* F0, const R* F1, const R* rtrec, R* esc,
                        double* slab, int E, int tid, int lane, int wave) {
  for (int rg = 0; rg < p.nranges; ++rg) {
    if constexpr (SK != kAny) {
      rt_range<CH, R, GEO, SK, CUT>(p, p.tc[0], F0, rtrec, esc, slab, E, tid, lane, wave, rg);
    } else {
      const int set = p.range_kind[rg] == 1 ? 0 : 1;
      const int kind = set ? p.sca_lo.kind : p.sca.kind;
      const R* F = set ? F1 : F0;
      if (kind == kSpl) {
        // a plane's Spline pass of 8 bins spills ~40 registers per event: 4
        // bins per pass there (1-2% faster)
        constexpr int SCH = (GEO == 0 && CH > 4) ? 4 : CH;
        rt_range<SCH, R, GEO, kSpl, CUT>(p, p.tc[set], F, rtrec, esc, slab, E, tid, lane, wave, rg);
      } else if constexpr (GEO == 0) {
        if (kind == kMie)
          rt_range<CH, R, GEO, kMie, CUT>(p, p.tc[set], F, rtrec, esc, slab, E, tid, lane, wave, rg);
        else
          rt_range<CH, R, GEO, kRay, CUT>(p, p.tc[set], F, rtrec, esc, slab, E, tid, lane, wave, rg);
      }
    }
  }
}

// CB bins × events (cpu.jl:83-158): thread tid owns CB bins tid, tid + kThreads, ….
// CB bin: kf[0..2], Eh[3..4], Eh × kf [5..7], cos2ψ_b [8], sin2ψ_b [9], 1/kf_z [10].
// Work items are (bin, group) pairs, item = lb + nb·g: the events of a bin are
// split into cb_groups interleaved groups so that the items spread evenly over
// the workgroup (808 bins would leave 216 of 512 threads with half the work).
// A private slab holds one copy of the CB bins per group; shared slabs take
// every group with atomics.
// CB record: [0..7] the scaled forward and reversed fields (n > 1) or I and
// [32..33] J (n = 1); [8..15] frame matrix of the event direction, [16..23]
// of -k₂; [24..26] K·x, [27..29] K·x₁; [30], [31] K·τ of the layer at x and x₁
// (plane) or K²·(τ² - x·x) and K²·(τ² - x₁·x₁), both ≥ 0 (sphere); [39] kind
// (0 none, 1 n = 1, 2 n > 1); K = ExpK<R>::k, so that distances come out
// scaled for exp2 as in the peel-off.  CUT = false when no CB distance can
// exceed τ₀ (a sphere with 4τ < τ₀: τ_a ≤ 2τ and |τ_b + τ_ab| ≤ 4τ), where both
// cut-off tests (cpu.jl:97, 104) always pass.
template <typename R, int GEO, int SK, bool ATOMIC, bool CUT>
MCRT_DEV void cb_pass(const CoopParams<R>& p, const R* F, const R* S, const R* cbrec,
                      const double* cbdx, double* slab, int E, int tid) {
  const TabC<R>& t = p.tc[0];
  const int nb = p.Ntb * p.Npb;
  // split only iterations with enough events to share (uniform over the workgroup)
  const int G = E >= 8 * p.cb_groups ? p.cb_groups : 1;
  const R tau0k = p.tau0k;
  const int total = nb * G;
  // a work item's bin constants; the next item's are loaded before the flush
  // of the current one (loads issued after its atomics would wait for them)
  struct Item { R b[kBinW]; double kik[3]; };
  auto load = [&](int it, Item& q) {
    const int lb = it % nb;
    const R* src = p.cb_bin + kBinStride * lb;
#pragma unroll
    for (int c = 0; c < kBinW; ++c) q.b[c] = src[c];
    const double* kk = p.cb_kik + 4 * lb;
    q.kik[0] = kk[0]; q.kik[1] = kk[1]; q.kik[2] = kk[2];
  };
  Item cur;
  if (tid < total) load(tid, cur);
  for (int it = tid; it < total;) {
    const int lb = it % nb, grp = it / nb;
    const R* b = cur.b;
    const R kfx = b[0], kfy = b[1], kfz = b[2], ikz = b[10];
    const double kikx = cur.kik[0], kiky = cur.kik[1], kikz = cur.kik[2];
    R acc[4] = {R(0), R(0), R(0), R(0)};
    for (int e = grp; e < E; e += G) {
      const R* cr = cbrec + kCbRec * e;
      const int kind = kind_of(cr[39]);
      if (kind == 0) continue;
      const V3<R> xa = {cr[24], cr[25], cr[26]};
      R ta;  // K·τ_a
      if (GEO == 0) ta = kfz > R(0) ? -xa.z * ikz : (xa.z + cr[30]) * -ikz;
      else {
        R kx = kfx * xa.x + kfy * xa.y + kfz * xa.z;
        ta = bsqrt(kx * kx + cr[30]) - kx;
      }
      if (CUT && !(ta <= tau0k)) continue;
      const Plane<R> qa = scat_plane(cr + 8, kfx, kfy, kfz, b[3], b[4], b[5], b[6], b[7]);
      // the bin's Stokes vector o before the final rotation, and the factors
      // s0 (I, V) and s12 (Q, U) that scale it into the accumulation
      R o[4], s0, s12;
      if (kind == 1) {  // n = 1 (cpu.jl:131-158)
        peel_I<R, SK, true>(t, F, cr, cr + 32, qa, o);
        s0 = ExpK<R>::exp_neg(ta);
        s12 = s0 * qa.inv;
      } else {          // n > 1 (cpu.jl:83-127)
        const V3<R> xb = {cr[27], cr[28], cr[29]};
        Cx<R> a1 = {cr[0], cr[1]}, a2 = {cr[2], cr[3]};
        Cx<R> b1 = {cr[4], cr[5]}, b2 = {cr[6], cr[7]};
        const Plane<R> qb = scat_plane(cr + 16, kfx, kfy, kfz, b[3], b[4], b[5], b[6], b[7]);
        const double* dx = cbdx + kCbDx * e;
        const R tab = (R)dx[3];  // K·τ_ab
        R tb;
        if (GEO == 0) tb = kfz > R(0) ? -xb.z * ikz : (xb.z + cr[31]) * -ikz;
        else {
          R kx = kfx * xb.x + kfy * xb.y + kfz * xb.z;
          tb = bsqrt(kx * kx + cr[31]) - kx;
        }
        const bool bok = !CUT || M<R>::fabs(tb + tab) <= tau0k;
        // the attenuations e^{-τ/2} and the 1/ρ² of each plane ride on peel_E's ψ rotation
        const R na = ExpK<R>::exp_neg(R(0.5) * ta) * qa.inv;
        const R nb_ = bok ? ExpK<R>::exp_neg(R(0.5) * (tb + tab)) * qb.inv : R(0);
        peel_E<R, SK, true>(t, S, qa, na, a1, a2);
        peel_E<R, SK, true>(t, S, qb, nb_, b1, b2);
        // the phase kl·(k_i + k_f)·(x_a - x_b) (cpu.jl:94), kl and the unit folded
        // into the per-bin k_i + k_f
        const double dphi = kikx * dx[0] + kiky * dx[1] + kikz * dx[2];
        // only the intensities of the two fields are needed (cpu.jl:100, 108)
        const R Ia0 = abs2(a1) + abs2(a2);
        const R Ib0 = abs2(b1) + abs2(b2);
        R cph, sph_;
        phase_cs(dphi, cph, sph_);
        Cx<R> ep = {cph, sph_};
        Cx<R> e1 = a1 * ep + b1, e2 = a2 * ep + b2;
        I_from_E(e1, e2, o);
        // Ia/(Ia + Ib) (cpu.jl:114); 0 where both fields vanish (0·finite)
        s0 = Ia0 * recip(fmax_r(Ia0 + Ib0, std::numeric_limits<R>::min()));
        s12 = s0;
      }
      acc[0] = fma_r(o[0], s0, acc[0]);
      acc[1] = fma_r(o[1], s12, acc[1]);
      acc[2] = fma_r(o[2], s12, acc[2]);
      acc[3] = fma_r(o[3], s0, acc[3]);
    }
    // the final rotation (cpu.jl:118-124) is linear in (Q, U) with the bin's own
    // coefficients, and Q and U share the scale s12: it acts once on the sums
    rotate(acc, b[8], b[9]);
    const int nit = it + kThreads;
    Item nxt;
    if (nit < total) load(nit, nxt);
    const int64_t ob = 5 + p.nrt + (ATOMIC ? 0 : (int64_t)grp * p.ncb) + 4 * (int64_t)lb;
#pragma unroll
    for (int c = 0; c < 4; ++c) unsafeAtomicAdd(slab + ob + c, (double)acc[c]);
    if (nit >= total) break;
    it = nit;
    cur = nxt;
  }
}

// LDS blocks of one workgroup for the solve p, as byte offsets from the
// dynamic LDS base (kNoLds: absent); every block is a multiple of 16 bytes and
// the host sizes the launch with the same function.  The event records are
// sized by the solve's active slots.  Set 0's F first: its LDS address is then
// the lookup offset itself.
constexpr size_t kNoLds = ~size_t(0);
struct LdsOff { size_t F, rtrec, cbrec, cbdx, esc, wcnt, C, S, F1, end; };
template <typename R, int SK, bool CB>
MCRT_HD LdsOff lds_offsets(const CoopParams<R>& p) {
  LdsOff o;
  size_t q = 0;
  auto take = [&q](size_t bytes) { const size_t r = q; q += (bytes + 15) & ~size_t(15); return r; };
  const int k0 = SK == kAny ? p.sca.kind : SK;
  const bool tab0 = k0 != kRay;
  const size_t n = (size_t)p.slots;
  o.F = tab0 ? take(sizeof(R) * f_stride(k0) * (size_t)(p.sca.Nt + 1)) : kNoLds;
  o.rtrec = take(sizeof(R) * kRtRec * n);
  o.cbrec = CB ? take(sizeof(R) * kCbRec * n) : kNoLds;
  o.cbdx = CB ? take(sizeof(double) * kCbDx * n) : kNoLds;
  o.esc = take(sizeof(R) * n * kWaves * 2);
  o.wcnt = take(sizeof(int) * 32);
  // csrn of the transport's sampling (kAny samples from the tables in HBM)
  o.C = (SK != kAny && tab0) ? take(sizeof(R) * (size_t)(p.sca.rp + 1)) : kNoLds;
  o.S = (CB && tab0 && p.stage_s) ? take(sizeof(R) * 4 * (size_t)(p.sca.Nt + 1)) : kNoLds;
  o.F1 = (SK == kAny && !p.lo_shared && p.sca_lo.kind != kRay)
             ? take(sizeof(R) * f_stride(p.sca_lo.kind) * (size_t)(p.sca_lo.Nt + 1))
             : kNoLds;
  o.end = q;
  return o;
}

template <typename R> struct Lds {
  R* F;       // set 0's F table
  R* rtrec;
  R* cbrec;
  double* cbdx;
  R* esc;
  int* wcnt;
  R* C;       // set 0's csrn
  R* S;       // set 0's S table (NULL: read from HBM)
  R* F1;      // set 1's F table when it is not set 0's
};
template <typename R, int SK, bool CB>
MCRT_DEV Lds<R> lds_layout(const CoopParams<R>& p, unsigned char* q) {
  const LdsOff o = lds_offsets<R, SK, CB>(p);
  auto at = [q](size_t off) { return off == kNoLds ? nullptr : q + off; };
  Lds<R> L;
  L.F = (R*)at(o.F);
  L.rtrec = (R*)at(o.rtrec);
  L.cbrec = (R*)at(o.cbrec);
  L.cbdx = (double*)at(o.cbdx);
  L.esc = (R*)at(o.esc);
  L.wcnt = (int*)at(o.wcnt);
  L.C = (R*)at(o.C);
  L.S = (R*)at(o.S);
  L.F1 = (R*)at(o.F1);
  return L;
}

// Copy the solve's tables into LDS (the caller synchronises afterwards).
template <typename R>
MCRT_DEV void stage_tables(const CoopParams<R>& p, const Lds<R>& L, int tid) {
  if (L.F)
    for (int i = tid; i < f_stride(p.sca.kind) * (p.sca.Nt + 1); i += kThreads) L.F[i] = p.sca.F[i];
  if (L.C)
    for (int i = tid; i <= p.sca.rp; i += kThreads) L.C[i] = p.sca.csrn[i];
  if (L.S)
    for (int i = tid; i < 4 * (p.sca.Nt + 1); i += kThreads) L.S[i] = p.sca.S[i];
  if (L.F1)
    for (int i = tid; i < f_stride(p.sca_lo.kind) * (p.sca_lo.Nt + 1); i += kThreads)
      L.F1[i] = p.sca_lo.F[i];
}

// The CB pass with set 0's tables (cpu.jl:50 takes layer 1's scatterer);
// kAny dispatches on its kind.
template <typename R, int GEO, int SK, bool ATOMIC>
MCRT_DEV void cb_all(const CoopParams<R>& p, const R* F, const R* S, const R* cbrec,
                     const double* cbdx, double* slab, int E, int tid) {
  if constexpr (SK != kAny) {
    if (GEO == 1 && !p.cb_cut) cb_pass<R, GEO, SK, ATOMIC, false>(p, F, S, cbrec, cbdx, slab, E, tid);
    else cb_pass<R, GEO, SK, ATOMIC, true>(p, F, S, cbrec, cbdx, slab, E, tid);
  } else {
    if (p.sca.kind == kSpl) {
      cb_all<R, GEO, kSpl, ATOMIC>(p, F, S, cbrec, cbdx, slab, E, tid);
    } else if constexpr (GEO == 0) {
      if (p.sca.kind == kMie) cb_all<R, GEO, kMie, ATOMIC>(p, F, S, cbrec, cbdx, slab, E, tid);
      else cb_all<R, GEO, kRay, ATOMIC>(p, F, S, cbrec, cbdx, slab, E, tid);
    }
  }
}

// The CB path state and record of one slot's event: the slot sp holds the
// event's direction k (arriving at x), position x, layer idx, Stokes vector I
// after absorption and event number n (the transport's scatter has written
// only kp, xp, idxp, Ip beside them); rec is its RT record (near flags).
//  * forward_E advanced one step (utils.jl:205-213) and the reversed path of
//    backward_E as a normalised running 2×2 product (utils.jl:230-238);
//  * collect_cb's set-up for n > 1 (cpu.jl:53-81) or the n = 1 record
//    (cpu.jl:131-158), or an empty record (kind 0).
// CB record: [0..7] the scaled forward and reversed fields (n > 1) or I and
// [32..33] J (n = 1); [8..15] frame matrix of the event direction, [16..23]
// of -k₂; [24..26] K·x, [27..29] K·x₁; [30], [31] K·τ of the layer at x and x₁
// (plane) or K²·(τ² - x·x) and K²·(τ² - x₁·x₁) (sphere); [39] kind;
// dx: x - x₁ and K·τab in FP64.
template <typename R, int GEO>
MCRT_DEV void cb_event(const CoopParams<R>& p, Slot<R>* sp, const R* rec, R* cr, double* dx) {
  const int64_t n = sp->n;
  const V3<R> x = sp->x;
  const int idx = sp->idx;
  if (n == 1) {
    sp->x1d = {(double)x.x, (double)x.y, (double)x.z};
    sp->idx1 = idx;
  }
  const V3<R> k = sp->k;
  const V3<double> kd = {(double)k.x, (double)k.y, (double)k.z};
  Cx<double> Ea1 = sp->Ea1, Ea2 = sp->Ea2;
  double enorma = sp->enorma;
  if (n >= 2) {  // forward_E advanced by one step (utils.jl:205-213)
    update_E(p.sca64, Ea1, Ea2, kd, sp->ksf);
    double a = hypot(Ea1.re, Ea1.im), b = hypot(Ea2.re, Ea2.im);
    double norm2 = sqrt(a * a + b * b);
    double r = p.norm0 / norm2;
    Ea1 = Ea1 * r;
    Ea2 = Ea2 * r;
    enorma += log(p.norm0 / norm2);
    sp->Ea1 = Ea1;
    sp->Ea2 = Ea2;
    sp->enorma = enorma;
    sp->ksf = sph(kd);
  }
  // reversed path as a running product (utils.jl:230-238)
  Lmap<double> Q = {{1, 0}, {0, 0}, {0, 0}, {1, 0}};  // (used for n > 1 only)
  double lq = 0.0;
  Fr<double> kbs = {1, 0, 1, 0};
  if (n == 2) {
    Q = {{1, 0}, {0, 0}, {0, 0}, {1, 0}};
    lq = 0.0;
    kbs = sph(-kd);
    sp->Q = Q;
    sp->lq = lq;
    sp->kbs = kbs;
  } else if (n >= 3) {
    Lmap<double> L = update_E_map(p.sca64, -sp->kprev, sph(-kd));
    Q = mul(sp->Q, L);
    lq = sp->lq;
    lmap_normalise(Q, lq);
    kbs = sp->kbs;
    sp->Q = Q;
    sp->lq = lq;
  }
  sp->kprev = kd;

  const int cb_on = p.compute_cb && (rec[18] != R(0) || rec[19] != R(0));
  const Fr<double> kas = sph(kd);
  const V3<double> xa = {(double)x.x, (double)x.y, (double)x.z};
  const V3<double> kf0 = {p.kf0[0], p.kf0[1], p.kf0[2]};
  const R K = ExpK<R>::k;
  if (cb_on && n > 1) {
    // collect_cb n > 1 set-up (cpu.jl:53-81) from the O(1) state
    const V3<double> x1d = sp->x1d;
    const int idx1 = sp->idx1;
    double ta = dist64(p, xa, kf0, idx);
    double tb = dist64(p, x1d, kf0, idx1);
    const R* I = sp->I;
    double Id[4] = {(double)I[0], (double)I[1], (double)I[2], (double)I[3]};
    double Iaf[4];
    update_I(p.sca64, Id, kf0, kas, Iaf);
    // reversed path: Q · L(-k_n | frame₀) · E_i, scaled by exp(lq + enorma)
    Fr<double> f0 = {p.f0[0], p.f0[1], p.f0[2], p.f0[3]};
    Lmap<double> L0 = update_E_map(p.sca64, -kd, f0);
    Cx<double> Ei1 = {p.Ei[0], p.Ei[1]}, Ei2 = {p.Ei[2], p.Ei[3]};
    Cx<double> u1 = L0.a * Ei1 + L0.b * Ei2, u2 = L0.c * Ei1 + L0.d * Ei2;
    Cx<double> v1 = Q.a * u1 + Q.b * u2, v2 = Q.c * u1 + Q.d * u2;
    double sc = exp(lq + enorma);
    Cx<double> Eb1 = v1 * sc, Eb2 = v2 * sc;
    Cx<double> Eaf1 = Ea1, Eaf2 = Ea2;
    update_E(p.sca64, Eaf1, Eaf2, kf0, kas);
    // raa (cpu.jl:68-69); zero where rounding leaves Iaf₁ ≤ 0 or the
    // field is zero (the reference would raise DomainError or give NaN)
    const double eaf2 = Eaf1.re * Eaf1.re + Eaf1.im * Eaf1.im + Eaf2.re * Eaf2.re +
                        Eaf2.im * Eaf2.im;
    const double raa = (Iaf[0] > 0 && eaf2 > 0) ? sqrt(Iaf[0] / eaf2) : 0.0;
    cr[0] = (R)(Ea1.re * raa); cr[1] = (R)(Ea1.im * raa);
    cr[2] = (R)(Ea2.re * raa); cr[3] = (R)(Ea2.im * raa);
    cr[4] = (R)(Eb1.re * raa); cr[5] = (R)(Eb1.im * raa);
    cr[6] = (R)(Eb2.re * raa); cr[7] = (R)(Eb2.im * raa);
    R ma_[8], mb_[8];
    frame_matrix(Fr<R>{(R)kas.ct, (R)kas.st, (R)kas.cp, (R)kas.sp}, ma_);
    frame_matrix(Fr<R>{(R)kbs.ct, (R)kbs.st, (R)kbs.cp, (R)kbs.sp}, mb_);
    for (int c = 0; c < 8; ++c) { cr[8 + c] = ma_[c]; cr[16 + c] = mb_[c]; }
    // positions and path lengths scaled by K for exp2 (see cb_pass)
    const V3<R> xb = {(R)x1d.x, (R)x1d.y, (R)x1d.z};
    cr[24] = K * x.x; cr[25] = K * x.y; cr[26] = K * x.z;
    cr[27] = K * xb.x; cr[28] = K * xb.y; cr[29] = K * xb.z;
    if (GEO == 0) {
      cr[30] = p.isinf[idx] ? R(INFINITY) : K * p.tau[idx];
      cr[31] = p.isinf[idx1] ? R(INFINITY) : K * p.tau[idx1];
    } else {  // τ² - x·x of the sphere's distance (sphere.jl:69-73), ≥ 0
      cr[30] = K * K * fmax_r(p.tau[0] * p.tau[0] - dot(x, x), R(0));
      cr[31] = K * K * fmax_r(p.tau[0] * p.tau[0] - dot(xb, xb), R(0));
    }
    dx[0] = xa.x - x1d.x; dx[1] = xa.y - x1d.y; dx[2] = xa.z - x1d.z;
    dx[3] = (ta - tb) * (double)ExpK<R>::k;  // K·τab
    cr[39] = kind_word<R>(2);
  } else if (cb_on) {
    // collect_cb n = 1 (cpu.jl:131-158)
    R ma_[8];
    frame_matrix(Fr<R>{(R)kas.ct, (R)kas.st, (R)kas.cp, (R)kas.sp}, ma_);
    for (int c = 0; c < 8; ++c) cr[8 + c] = ma_[c];
    const R* I = sp->I;
    cr[0] = I[0]; cr[1] = I[1]; cr[2] = I[2]; cr[3] = I[3];
    cr[32] = R(2) * I[1]; cr[33] = R(-2) * I[2];
    cr[24] = K * x.x; cr[25] = K * x.y; cr[26] = K * x.z;
    cr[30] = GEO == 0 ? (p.isinf[idx] ? R(INFINITY) : K * p.tau[idx])
                      : K * K * fmax_r(p.tau[0] * p.tau[0] - dot(x, x), R(0));
    cr[39] = kind_word<R>(1);
  } else {
    cr[39] = kind_word<R>(0);
  }
}

// Rays [rb, re) of the solve p through one workgroup, with the tables staged
// in L.  Bins are flushed after every iteration into the slab by FP64 atomic
// adds: sums private to the workgroup in slab[0, nout) (slab[nout, 2·nout)
// holds compensations, written only for the energy budget and zero
// elsewhere), or with ATOMIC a slab shared with other chunks.  The energy
// budget goes to slab[0..4] at the end and the work counters to p.counters.
// slots: this workgroup's SLOTS ray-state records (empty again on return).
// With preempt (a host flag), the run stops at an iteration boundary once the
// flag is set, after at least one iteration: the rays in flight stay in slots
// (with their live flags), *state receives the first ray not yet started, and
// a later call with resume continues them from rb = that ray; a run that ends
// stores -2 there.  The iterations, and so the per-iteration bin sums, are
// the same as those of an uninterrupted run.
template <typename R, int GEO, int SK, bool CB, bool ATOMIC>
MCRT_DEV void run_rays(const CoopParams<R>& p, const Lds<R>& L, int64_t rb, int64_t re,
                       double* slab, Slot<R>* slots, const unsigned int* preempt = nullptr,
                       bool resume = false, int64_t* state = nullptr) {
  constexpr int SLOTS = Shape<R>::kSlots;
  constexpr int SW = SLOTS / 64;  // waves owning slots
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const bool owner = tid < p.slots;
  R* const rtrec = L.rtrec;
  R* const cbrec = L.cbrec;
  double* const cbdx = L.cbdx;
  R* const esc = L.esc;
  int* const wcnt = L.wcnt;

  // the transport's view of set 0 (F/csrn in LDS; kAny reads each layer's
  // tables from HBM, lsca)
  ScaView<R> sca = p.sca;
  if constexpr (SK == kMie) {
    sca.F = L.F;
    sca.csrn = L.C;
  }
  const R* const Sbins = L.S ? (const R*)L.S : p.sca.S;
  const R* const F1 = p.lo_shared ? L.F : L.F1;  // set 1's F (kAny)
  const Geo<R> geo{&p};

  SlotAcc* const acc = (SlotAcc*)(slots + SLOTS) + (owner ? tid : 0);
  if (owner) {
#pragma unroll
    for (int c = 0; c < 5; ++c) acc->A[c] = 0.0;
#pragma unroll
    for (int c = 0; c < kCntN; ++c) acc->cnt[c] = 0ull;
  }
  double* const Aacc = acc->A;
  unsigned long long* const cnt = acc->cnt;
  double* slabc = slab + p.nout;
  const bool can_dt = (p.geo_kind == 1) || (p.nl == 1 && !p.isinf[0]);

  // Per-slot ray state lives in HBM between the transport and update phases
  // so that the bin loops run with a small register footprint.
  Slot<R>* const myslot = slots + (owner ? tid : 0);
  int live = (resume && owner) ? myslot->live : 0;  // this slot holds a ray
  int64_t next_ray = rb;  // identical in every thread
  int e_me = 0;
  bool stopped = false;

  for (int iter = 0;; ++iter) {
    // wcnt[30]: the preempt flag as read during the previous iteration
    if (preempt && iter > 0 && wcnt[30]) {
      stopped = true;
      break;
    }
    {
      // value-initialised: a slot that starts the iteration empty gets every
      // field it later reads from the refill, and no field of s is left for
      // the optimiser to choose (uninitialised, the FP64 general-kind kernel
      // was compiled to start a refilled ray's transport at RNG draw 0)
      RaySt<R> s{};
      if (live) s = *myslot;
      s.live = live;
      // ------------------------------------- refill free slots (in order)
      {
        unsigned long long need = __ballot(owner && !s.live);
        if (lane == 0) wcnt[wave] = __popcll(need);
        __syncthreads();
        int off = 0, tot = 0;
        for (int w = 0; w < SW; ++w) {
          if (w < wave) off += wcnt[w];
          tot += wcnt[w];
        }
        off += __popcll(need & ((1ull << lane) - 1ull));
        if (owner && !s.live && next_ray + off < re) {
          // a new ray: initial_x, direct transmission, initial_propagation
          s.live = 1;
          s.ray = p.ray0 + (uint64_t)(next_ray + off);
          Rng rng{p.seed, s.ray, p.instance, 0};
#pragma unroll
          for (int c = 0; c < 4; ++c) s.I[c] = p.I0[c];
          s.k = {p.k0[0], p.k0[1], p.k0[2]};
          if (GEO == 0) {
            s.x = {R(0), R(0), R(0)};
          } else {  // sphere.jl:43-47
            R tau = p.tau[0];
            R r = tau * M<R>::sqrt(rng.template u<R>());
            R ph = M<R>::two_pi * rng.template u<R>();
            R sp_, cp_;
            M<R>::sincos(ph, &sp_, &cp_);
            s.x = {r * cp_, r * sp_, M<R>::sqrt(tau * tau - r * r)};
          }
          s.idx = 0;
          s.n = 0;
          if (can_dt) {  // cpu.jl:175-180
            R tp = geo.dist(s.x, s.k, 0);
            R dt = tp <= p.tau0 ? M<R>::exp(-tp) : R(0);
            Aacc[0] += (double)(s.I[0] * dt);
#pragma unroll
            for (int c = 0; c < 4; ++c) s.I[c] = s.I[c] * (R(1) - dt);
          }
          {  // initial_propagation (plane.jl:77-95, sphere.jl:54-67)
            R t;
            if (GEO == 0) {
              if (p.isinf[0]) {
                t = -M<R>::log(rng.template u<R>());
              } else {
                R tp = geo.dist(s.x, s.k, 0);
                if (can_dt && tp <= p.tau0) {
                  t = -M<R>::log(R(1) - rng.template u<R>() * (R(1) - M<R>::exp(-tp)));
                } else {
                  t = tp + R(1);
                  for (int it = 0; t > tp && it < (1 << 24); ++it) t = -M<R>::log(rng.template u<R>());
                }
              }
            } else {
              R tp = R(2) * s.x.z;
              if (tp <= p.tau0) {
                t = -M<R>::log(R(1) - rng.template u<R>() * (R(1) - M<R>::exp(-tp)));
              } else {
                t = tp + R(1);
                for (int it = 0; t > tp && it < (1 << 24); ++it)
                  t = -M<R>::log(R(1) - rng.template u<R>());
              }
            }
            s.x = {s.x.x + t * s.k.x, s.x.y + t * s.k.y, s.x.z + t * s.k.z};
          }
          s.draw = rng.draw;
          ++cnt[kCntRays];
          if (!(s.I[0] >= p.min_int && p.maxscat > 0)) {  // the loop of cpu.jl:186 never runs
            if (!isfinite((double)s.I[0])) ++cnt[kCntNonfinite];
            Aacc[4] += (double)s.I[0];
            s.live = 0;
          }
          if (CB) {
            CbSt& cs = *myslot;
            cs.Ea1 = {p.Ei[0], p.Ei[1]};
            cs.Ea2 = {p.Ei[2], p.Ei[3]};
            cs.enorma = 0.0;
            cs.ksf = {p.f0[0], p.f0[1], p.f0[2], p.f0[3]};
            cs.lq = 0.0;
          }
        }
        next_ray += tot;
      }

      // --------------------------------- compaction of the live slots
      unsigned long long lv = __ballot(owner && s.live);
      if (lane == 0) wcnt[16 + wave] = __popcll(lv);
      __syncthreads();
      int E = 0, eoff = 0;
      for (int w = 0; w < SW; ++w) {
        if (w < wave) eoff += wcnt[16 + w];
        E += wcnt[16 + w];
      }
      if (E == 0) break;
      e_me = eoff + __popcll(lv & ((1ull << lane) - 1ull));

      // ------------------------------------ transport: one event per slot
      if (owner && s.live) {
        Rng rng{p.seed, s.ray, p.instance, s.draw};
        s.n += 1;
        // scatter + free path with rejection (utils.jl:294-323, cpu.jl:194-235)
        // with the scatterer of the ray's layer (scatterer(g, idx))
        {
          const ScaView<R> lsc = SK == kAny ? p.lsca[s.idx] : sca;
          Fr<R> ks = sph(s.k);
          for (int it = 0;; ++it) {
            R ct = phase_sample(lsc, rng.template u<R>());
            R P[6];
            phase_eval(lsc, ct, P);
            R eq = -P[1] * s.I[1] / (P[0] * s.I[0]);
            R eu = -P[1] * s.I[2] / (P[0] * s.I[0]);
            R ee = M<R>::sqrt(eq * eq + eu * eu);
            R phi;
            if (ee > R(1e-12)) {
              R gam = M<R>::acos(clamp1(eq / ee));
              if (eu < R(0)) gam = M<R>::two_pi - gam;
              R eec = ee < R(1) ? ee : R(1);
              R ma = M<R>::four_pi * rng.template u<R>() + gam - eec * M<R>::sin(gam);
              R ea = kepler(ma, eec);
              phi = R(0.5) * (ea - gam);
            } else {
              phi = M<R>::two_pi * rng.template u<R>();
            }
            R st = sqrt0(R(1) - ct * ct);
            R sph_, cph_;
            M<R>::sincos(phi, &sph_, &cph_);
            s.kp = ray_to_norm(V3<R>{st * cph_, st * sph_, ct}, ks);
            R Ipt[4];
            update_I(lsc, s.I, s.kp, ks, Ipt);
            R rr = s.I[0] / Ipt[0];
#pragma unroll
            for (int c = 0; c < 4; ++c) s.Ip[c] = Ipt[c] * rr;
            R t = -M<R>::log(rng.template u<R>());
            s.xp = {s.x.x + t * s.kp.x, s.x.y + t * s.kp.y, s.x.z + t * s.kp.z};
            int pb;
            if (GEO == 0)
              pb = s.xp.z > R(0) ? -1 : ((!p.isinf[s.idx] && s.xp.z < -p.tau[s.idx]) ? 1 : 0);
            else
              pb = M<R>::sqrt(dot(s.xp, s.xp)) > p.tau[0] ? -2 : 0;
            s.idxp = s.idx + pb;
            bool ok;
            if (pb == -2) ok = false;
            else if (pb == -1) {
              ok = s.idxp >= 0;
              if (ok) s.xp = {s.x.x, s.x.y, -p.tau[s.idx]};
            } else if (pb == 1) {
              ok = s.idxp <= p.nl - 1;
              if (ok) s.xp = {s.x.x, s.x.y, R(0)};
            } else ok = true;
            if (ok) break;
            ++cnt[kCntRej];
            if (it >= (1 << 24)) { s.idxp = s.idx; s.xp = s.x; break; }
          }
        }
        s.draw = rng.draw;
        // absorption (cpu.jl:238-239)
        R om = p.omega[s.idx];
        Aacc[1] += (double)((R(1) - om) * s.I[0]);
#pragma unroll
        for (int c = 0; c < 4; ++c) s.I[c] = s.I[c] * om;
        // near flags (cpu.jl:245-246)
        R dfirst, dlast;
        if (GEO == 0) {
          dfirst = s.idx == 0 ? M<R>::fabs(s.x.z) : R(INFINITY);
          dlast = (s.idx == p.nl - 1 && !p.isinf[s.idx]) ? p.tau[s.idx] - M<R>::fabs(s.x.z) : R(INFINITY);
        } else {
          dfirst = p.tau[0] - M<R>::sqrt(dot(s.x, s.x));
          dlast = R(INFINITY);
        }
        const int near_up = dfirst <= p.tau0;
        const int near_lo = dlast <= p.tau0;
        // RT event record: frame matrix, I, position, flags
        R* rec = rtrec + kRtRec * e_me;
        frame_matrix(sph(s.k), rec);
        rec[8] = s.I[0]; rec[9] = s.I[1]; rec[10] = s.I[2]; rec[11] = s.I[3];
        rec[12] = R(2) * s.I[1]; rec[13] = R(-2) * s.I[2];
        {
          const R K = ExpK<R>::k;
          rec[14] = K * s.x.x; rec[15] = K * s.x.y; rec[16] = K * s.x.z;
          if (GEO == 0) rec[17] = p.isinf[s.idx] ? R(INFINITY) : K * p.tau[s.idx];
          else rec[17] = K * K * fmax_r(p.tau[0] * p.tau[0] - dot(s.x, s.x), R(0));
        }
        rec[18] = R(near_up);
        rec[19] = R(near_lo);
        for (int w = 0; w < kWaves; ++w) {
          esc[(e_me * kWaves + w) * 2] = R(0);
          esc[(e_me * kWaves + w) * 2 + 1] = R(0);
        }
      }
      live = s.live;
      if (owner && live) static_cast<RaySt<R>&>(*myslot) = s;
      if (tid == 0) wcnt[31] = E;
    }
    __syncthreads();
    const int E = wcnt[31];

    // ----------------- CB path state and event records (cpu.jl:38-81), in a
    // phase of their own so that the FP64 path state is not live across the
    // transport's scatter loop
    if (CB) {
      if (owner && live)
        cb_event<R, GEO>(p, myslot, rtrec + kRtRec * e_me, cbrec + kCbRec * e_me,
                         cbdx + kCbDx * e_me);
      __syncthreads();
    }

    // ------------------------------- peel-off: bins × events (cpu.jl:6-31)
    constexpr int RCH =
        (ATOMIC && GEO == 0 && SK != kRay) ? Shape<R>::kChunkBatchPlane : Shape<R>::kChunk;
    if (GEO == 1 && !p.rt_cut)
      rt_ranges<RCH, R, GEO, SK, false>(p, L.F, F1, rtrec, esc, slab, E, tid, lane, wave);
    else
      rt_ranges<RCH, R, GEO, SK, true>(p, L.F, F1, rtrec, esc, slab, E, tid, lane, wave);

    // ------------------------------------------ CB bins × events (cpu.jl:38-160)
    if (CB && p.compute_cb) cb_all<R, GEO, SK, ATOMIC>(p, L.F, Sbins, cbrec, cbdx, slab, E, tid);
    __syncthreads();
    // read here, published by the barrier at the end of the iteration
    if (preempt && tid == 0)
      wcnt[30] = __hip_atomic_load(preempt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u;

    // ---------------------- weight update and termination (cpu.jl:263-272)
    if (owner && live) {
      RaySt<R> s = *myslot;
      R Iref = R(0), Itrans = R(0);
      for (int w = 0; w < kWaves; ++w) {
        Iref += esc[(e_me * kWaves + w) * 2];
        Itrans += esc[(e_me * kWaves + w) * 2 + 1];
      }
      const int fl = (rtrec[kRtRec * e_me + 18] != R(0)) | ((rtrec[kRtRec * e_me + 19] != R(0)) << 1);
      for (int rg = 0; rg < p.nranges; ++rg)
        if (fl & (p.range_kind[rg] == 1 ? 1 : 2))
          cnt[kCntRt] += (unsigned long long)(p.range_e[rg] - p.range_b[rg]) * p.Np;
      if (CB && p.compute_cb && (fl & 3)) {
        cnt[kCntCb] += (unsigned long long)p.Ntb * p.Npb;
        if (s.n == 1) cnt[kCntCb1] += (unsigned long long)p.Ntb * p.Npb;
      }
      Aacc[2] += (double)Iref;
      Aacc[3] += (double)Itrans;
      R Iesc = Iref + Itrans;
      R sc = s.I[0] - Iesc;
#pragma unroll
      for (int c = 0; c < 4; ++c) s.I[c] = s.Ip[c] * sc / s.Ip[0];
      s.k = s.kp; s.x = s.xp; s.idx = s.idxp;
      ++cnt[kCntEvents];
      if (!(s.I[0] >= p.min_int && s.n < p.maxscat)) {
        if (!isfinite((double)s.I[0])) ++cnt[kCntNonfinite];
        Aacc[4] += (double)s.I[0];
        live = 0;
      }
      if (live) static_cast<RaySt<R>&>(*myslot) = s;
    }
    __syncthreads();
  }

  if (stopped && owner) myslot->live = live;  // the other fields are current

  // ---------------------------------------- energy budget and counters
  {
    // reuse the record area: 5·slots doubles (it holds kRtRec·slots R words)
    double* buf = (double*)rtrec;
    const int ns = p.slots;
    if (owner)
      for (int c = 0; c < 5; ++c) buf[c * ns + tid] = Aacc[c];
    __syncthreads();
    if (tid < 5) {
      double sum = 0.0, cmp = 0.0;
      for (int i = 0; i < ns; ++i) kahan(&sum, &cmp, buf[tid * ns + i]);
      if (ATOMIC) {
        unsafeAtomicAdd(slab + tid, sum - cmp);
      } else {
        slab[tid] = sum;
        slabc[tid] = cmp;
      }
    }
  }
  if (owner)
    for (int c = 0; c < kCntN; ++c)
      if (cnt[c]) atomicAdd(p.counters + c, cnt[c]);
  // next_ray advances by the free slots of each refill, so it may pass re
  if (state && tid == 0) *state = stopped ? (next_ray < re ? next_ray : re) : int64_t(-2);
}

// One solve: workgroup g transports rays [N·g/G, N·(g+1)/G) into slab g.
template <typename R, int GEO, int SK, bool CB>
__global__ __launch_bounds__(kThreads) void coop_kernel(const CoopParams<R>* __restrict__ pp) {
  const CoopParams<R>& p = *pp;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const Lds<R> L = lds_layout<R, SK, CB>(p, lds_raw);
  stage_tables<R>(p, L, threadIdx.x);
  const int64_t G = gridDim.x;
  const int64_t rb = p.N * (int64_t)blockIdx.x / G;
  const int64_t re = p.N * ((int64_t)blockIdx.x + 1) / G;
  __syncthreads();
  run_rays<R, GEO, SK, CB, false>(p, L, rb, re, p.slabs + (size_t)blockIdx.x * 2 * p.nout,
                                  slot_block<R>(p.slot_mem, blockIdx.x));
}

// A uniform pointer in scalar registers (readfirstlane of both halves), so
// that a called function loads through it with scalar loads.
template <typename T> MCRT_DEV T* uniform_ptr(T* p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (T*)(((unsigned long long)hi << 32) | lo);
}

// One chunk of a batched launch, for one solve variant.  Not inlined: each
// variant is register-allocated on its own (inlined into the one persistent
// kernel, every variant's loops would pay for the register pressure of the
// heaviest).  It reaches LDS through the kernel's dynamic LDS symbol, so its
// LDS accesses stay LDS instructions.
extern __shared__ __align__(16) unsigned char batch_lds[];
template <typename R, int GEO, int SK, bool CB>
__device__ __attribute__((noinline)) void batch_chunk(const CoopParams<R>* pp_, int fresh,
                                                      const BatchChunk* chp_, Slot<R>* slots_,
                                                      const unsigned int* preempt_,
                                                      int64_t* state_) {
  const CoopParams<R>& p = *uniform_ptr(pp_);
  const BatchChunk* chp = uniform_ptr(chp_);
  Slot<R>* const slots = uniform_ptr(slots_);
  const unsigned int* preempt = uniform_ptr(preempt_);
  int64_t* const state = uniform_ptr(state_);
  const Lds<R> L = lds_layout<R, SK, CB>(p, batch_lds);
  if (__builtin_amdgcn_readfirstlane(fresh)) {
    stage_tables<R>(p, L, threadIdx.x);
    __syncthreads();
  }
  run_rays<R, GEO, SK, CB, true>(p, L, chp->rb, chp->re, chp->slab, slots, preempt,
                                 chp->resume != 0, state);
}

template <typename R, int GEO, int SK>
MCRT_DEV void batch_chunk_cb(const CoopParams<R>* pp, int fresh, const BatchChunk* chp,
                             Slot<R>* slots, const unsigned int* preempt, int64_t* state) {
  if (pp->compute_cb) batch_chunk<R, GEO, SK, true>(pp, fresh, chp, slots, preempt, state);
  else batch_chunk<R, GEO, SK, false>(pp, fresh, chp, slots, preempt, state);
}

template <typename R, int GEO>
MCRT_DEV void batch_chunk_geo(const CoopParams<R>* pp, int fresh, const BatchChunk* chp,
                              Slot<R>* slots, const unsigned int* preempt, int64_t* state) {
  if (pp->sk == kMie) batch_chunk_cb<R, GEO, kMie>(pp, fresh, chp, slots, preempt, state);
  else if (pp->sk == kRay) batch_chunk_cb<R, GEO, kRay>(pp, fresh, chp, slots, preempt, state);
  else batch_chunk_cb<R, GEO, kAny>(pp, fresh, chp, slots, preempt, state);
}

// Many solves in one persistent launch: each workgroup takes the next chunk
// (a ray range of one solve) from the queue until it is empty, restaging the
// tables when the solve changes.  A chunk's rays do not share iterations with
// any other chunk, so its per-iteration sums are the same whichever workgroup
// runs it and whatever else is in the launch; they are added into one of the
// solve's batch slabs with FP64 atomics.  With preempt, a workgroup claims
// its first chunk whatever the flag says (so that every launch advances) and
// no further one once the flag is set.
template <typename R>
__global__ __launch_bounds__(kThreads) void batch_kernel(const BatchChunk* __restrict__ chunks,
                                                         int nchunks, unsigned int* __restrict__ next,
                                                         void* slot_mem, int64_t* __restrict__ states,
                                                         const unsigned int* preempt) {
  __shared__ int chunk_id;
  Slot<R>* const own = slot_block<R>(slot_mem, blockIdx.x);
  const void* cur = nullptr;
  for (int claims = 0;; ++claims) {
    if (threadIdx.x == 0) {
      const bool stop = preempt && claims > 0 &&
                        __hip_atomic_load(preempt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u;
      chunk_id = stop ? nchunks : (int)atomicAdd(next, 1u);
    }
    __syncthreads();
    const int c = chunk_id;
    __syncthreads();
    if (c >= nchunks) break;
    const BatchChunk* const chp = chunks + c;
    const void* const params = chp->params;
    const CoopParams<R>* const pp = (const CoopParams<R>*)params;
    const int fresh = params != cur;
    cur = params;
    Slot<R>* const slots = chp->cont ? (Slot<R>*)chp->cont : own;
    int64_t* const st = states ? states + c : nullptr;
    if (pp->geo_kind == 0) batch_chunk_geo<R, 0>(pp, fresh, chp, slots, preempt, st);
    else batch_chunk_geo<R, 1>(pp, fresh, chp, slots, preempt, st);
  }
}

// Known-answer evaluation (mcrt_plan_eval): the building blocks above with
// table set 0 (kind SK) on n fixed records, converted to R on entry.
// UPDATE_I/UPDATE_E records carry the bin constants of k in words 11..15
// (Eh₁, Eh₂, Eh × k), as the bin tables do.
template <typename R, int SK>
__global__ void eval_kernel(const CoopParams<R>* __restrict__ pp, int what, int64_t n,
                            const double* __restrict__ in, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const CoopParams<R>& p = *pp;
  const TabC<R>& t = p.tc[0];
  // the transport's view of layer 1's scatterer
  const ScaView<R> tsc = p.sk == kAny ? p.lsca[0] : p.sca;
  const double* a = in + 16 * i;
  double* o = out + 4 * i;
  R r[4] = {R(0), R(0), R(0), R(0)};
  double d[4] = {0.0, 0.0, 0.0, 0.0};
  bool dbl = false;
  switch (what) {
    case MCRT_EVAL_UPDATE_I:
    case MCRT_EVAL_UPDATE_E: {
      R m[8];
      frame_matrix(Fr<R>{(R)a[0], (R)a[1], (R)a[2], (R)a[3]}, m);
      const Plane<R> q = scat_plane(m, (R)a[8], (R)a[9], (R)a[10], (R)a[11], (R)a[12], (R)a[13],
                                    (R)a[14], (R)a[15]);
      if (what == MCRT_EVAL_UPDATE_I) {
        const R I[4] = {(R)a[4], (R)a[5], (R)a[6], (R)a[7]};
        const R J[2] = {R(2) * I[1], R(-2) * I[2]};
        peel_I<R, SK>(t, p.sca.F, I, J, q, r);
        r[1] = r[1] * q.inv;
        r[2] = r[2] * q.inv;
      } else {
        Cx<R> E1 = {(R)a[4], (R)a[5]}, E2 = {(R)a[6], (R)a[7]};
        peel_E<R, SK, true>(t, p.sca.S, q, q.inv, E1, E2);
        r[0] = E1.re; r[1] = E1.im; r[2] = E2.re; r[3] = E2.im;
      }
      break;
    }
    case MCRT_EVAL_UPDATE_E_PATH: {
      Cx<double> E1 = {a[4], a[5]}, E2 = {a[6], a[7]};
      update_E(p.sca64, E1, E2, V3<double>{a[8], a[9], a[10]}, Fr<double>{a[0], a[1], a[2], a[3]});
      d[0] = E1.re; d[1] = E1.im; d[2] = E2.re; d[3] = E2.im;
      dbl = true;
      break;
    }
    case MCRT_EVAL_DISTANCE: {
      const int idx = (int)a[6];
      const Geo<R> geo{&p};
      d[0] = (double)geo.dist(V3<R>{(R)a[0], (R)a[1], (R)a[2]}, V3<R>{(R)a[3], (R)a[4], (R)a[5]}, idx);
      d[1] = dist64(p, V3<double>{a[0], a[1], a[2]}, V3<double>{a[3], a[4], a[5]}, idx);
      dbl = true;
      break;
    }
    case MCRT_EVAL_PHASE: {
      const Ph<R> P = phase_tab<R, SK>(t, p.sca.F, (R)a[0]);
      r[0] = P.p11; r[1] = P.p12; r[2] = P.p33; r[3] = P.p34;
      break;
    }
    case MCRT_EVAL_AMPLITUDE: {
      Cx<R> S11, S22;
      amp_tab<R, SK>(t, p.sca.S, (R)a[0], S11, S22);
      r[0] = S11.re; r[1] = S11.im; r[2] = S22.re; r[3] = S22.im;
      break;
    }
    case MCRT_EVAL_SAMPLE:
      r[0] = phase_sample(tsc, (R)a[0]);
      break;
    case MCRT_EVAL_PHASE_TRANSPORT: {
      R P[6];
      phase_eval(tsc, (R)a[0], P);
      r[0] = P[0]; r[1] = P[1]; r[2] = P[3]; r[3] = P[4];
      break;
    }
    case MCRT_EVAL_PHASE_DIAG: {
      const Ph<R> P = phase_tab<R, SK>(t, p.sca.F, (R)a[0]);
      R Q[6];
      phase_eval(tsc, (R)a[0], Q);
      r[0] = P.p22; r[1] = P.p44; r[2] = Q[2]; r[3] = Q[5];
      break;
    }
    case MCRT_EVAL_KEPLER:
      r[0] = kepler((R)a[0], (R)a[1]);
      break;
    case MCRT_EVAL_BIN_MATH:
      r[0] = ExpK<R>::exp_neg((R)(a[0] * ExpK<R>::k));
      r[1] = bsqrt((R)a[1]);
      r[2] = recip((R)a[2]);
      r[3] = t.scale > R(0) ? theta_index(t, (R)a[3]) / t.scale : R(0);
      break;
    case MCRT_EVAL_CB_PHASE:
      phase_cs(a[0], r[0], r[1]);
      break;
    default:
      break;
  }
  for (int c = 0; c < 4; ++c) o[c] = dbl ? d[c] : (double)r[c];
}

#if MCRT_PART(1)
// The value of output i in one private slab: output words A and Irt map
// one-to-one, a CB word is the sum over the slab's CB groups.
MCRT_DEV double slab_value(const double* sl, int64_t sn, int64_t i, int64_t nrt, int64_t ncb,
                           int groups) {
  if (i < 5 + nrt) return sl[i] - sl[sn + i];
  double v = 0.0;
  for (int q = 0; q < groups; ++q) {
    const int64_t j = i + (int64_t)q * ncb;
    v += sl[j] - sl[sn + j];
  }
  return v;
}

// out[i] = Σ_g slab_value_g(i) in workgroup order (Kahan); sn = slab words.
__global__ void slab_reduce(const double* __restrict__ slabs, int nslab, int64_t sn, int64_t nout,
                            double* __restrict__ A, double* __restrict__ Irt, double* __restrict__ Icb,
                            int64_t nrt, int64_t ncb, int groups) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nout) return;
  double s = 0.0, c = 0.0;
  for (int g = 0; g < nslab; ++g)
    kahan(&s, &c, slab_value(slabs + (size_t)g * 2 * sn, sn, i, nrt, ncb, groups));
  if (i < 5) A[i] = s;
  else if (i < 5 + nrt) Irt[i - 5] = s;
  else Icb[i - 5 - nrt] = s;
}

// Standard error of each normalised output T/N by batch means over the
// workgroup ray ranges: with S_g the raw sum of workgroup g over its n_g rays,
// v = Σ_g (S_g - n_g·T/N)² / n_g / (G - 1) estimates the per-ray variance and
// σ = √(v/N).  NaN when G < 2.
__global__ void slab_sigma(const double* __restrict__ slabs, int nslab, int64_t sn, int64_t nout,
                           int64_t N, int64_t nrt, int64_t ncb, int groups, double* __restrict__ sig) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nout) return;
  if (nslab < 2) { sig[i] = NAN; return; }
  double T = 0.0, c = 0.0;
  for (int g = 0; g < nslab; ++g)
    kahan(&T, &c, slab_value(slabs + (size_t)g * 2 * sn, sn, i, nrt, ncb, groups));
  const double mean = T / (double)N;
  double v = 0.0;
  for (int g = 0; g < nslab; ++g) {
    const int64_t ng = N * (g + 1) / nslab - N * g / nslab;  // the kernel's ray range
    const double d = slab_value(slabs + (size_t)g * 2 * sn, sn, i, nrt, ncb, groups) - (double)ng * mean;
    v += d * d / (double)ng;
  }
  sig[i] = sqrt(v / (double)(nslab - 1) / (double)N);
}

// Zero n 32-bit words (a kernel rather than a memset so that a captured graph
// holds no memset node).
__global__ void zero_words(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// Outputs of one batched solve from its batch slabs (slab b holds the raw sums
// of the chunks k ≡ b mod nbatch, brays[b] rays): out[i] = Σ_b S_b[i] in batch
// order, sig[i] the batch-means standard error of out[i]/N (NaN for one batch).
__global__ void batch_reduce(const double* __restrict__ slabs, int nbatch, int64_t nout, int64_t N,
                             const int64_t* __restrict__ brays, double* __restrict__ out,
                             double* __restrict__ sig) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nout) return;
  double T = 0.0, c = 0.0;
  for (int b = 0; b < nbatch; ++b) kahan(&T, &c, slabs[(size_t)b * nout + i]);
  T -= c;
  out[i] = T;
  if (nbatch < 2) {
    sig[i] = NAN;
    return;
  }
  const double mean = T / (double)N;
  double v = 0.0;
  for (int b = 0; b < nbatch; ++b) {
    const double nb = (double)brays[b];
    const double d = slabs[(size_t)b * nout + i] - nb * mean;
    v += d * d / nb;
  }
  sig[i] = sqrt(v / (double)(nbatch - 1) / (double)N);
}

// ------------------------------------------------------------------ host
// FP64 host restatements of the frame helpers used to precompute bin tables.
struct HFr { double ct, st, cp, sp; };
HFr h_sph(const double x[3]) {
  double s2 = x[0] * x[0] + x[1] * x[1];
  if (s2 < 1e-24) return {x[2] > 0 ? 1.0 : (x[2] < 0 ? -1.0 : x[2]), 0.0, 1.0, 0.0};
  double st = 1.0 - x[2] * x[2];
  double inv = 1.0 / std::sqrt(s2);
  return {x[2], std::sqrt(st < 0 ? 0 : st), x[0] * inv, x[1] * inv};
}
void h_ray_to_norm(const double x[3], HFr f, double o[3]) {
  o[0] = x[0] * f.ct * f.cp - x[1] * f.sp + x[2] * f.st * f.cp;
  o[1] = x[0] * f.ct * f.sp + x[1] * f.cp + x[2] * f.st * f.sp;
  o[2] = -x[0] * f.st + x[2] * f.ct;
}
void h_norm_to_ray(const double x[3], HFr f, double o[3]) {
  o[0] = x[0] * f.ct * f.cp + x[1] * f.ct * f.sp - x[2] * f.st;
  o[1] = -x[0] * f.sp + x[1] * f.cp;
  o[2] = x[0] * f.st * f.cp + x[1] * f.st * f.sp + x[2] * f.ct;
}
void h_get_basis(const double k[3], double Eh[3], double Ev[3]) {
  if (std::fabs(k[2]) > 1.0 - 2.220446049250313e-16) {
    Eh[0] = 0; Eh[1] = -1; Eh[2] = 0;
    Ev[0] = k[2] > 0 ? 1.0 : -1.0; Ev[1] = 0; Ev[2] = 0;
    return;
  }
  double t = 1.0 - k[2] * k[2];
  double ev3 = -std::sqrt(t < 0 ? 0 : t);
  double inv = 1.0 / ev3;
  double eh1 = k[1] * -inv, eh2 = k[0] * inv;
  Eh[0] = eh1; Eh[1] = eh2; Eh[2] = 0;
  Ev[0] = -k[2] * eh2; Ev[1] = -k[2] * eh1; Ev[2] = ev3;
}
// E_from_I (utils.jl:119-150)
void h_E_from_I(const double I[4], double E[4]) {
  const double tol = 1e-14, pi = 3.141592653589793;
  double A1, A2, ph;
  if (I[0] - std::fabs(I[1]) > tol * I[0]) {
    A1 = std::sqrt(0.5 * (I[0] - I[1]));
    A2 = std::sqrt(0.5 * (I[0] + I[1]));
    double q = std::sqrt(I[0] * I[0] - I[1] * I[1]);
    if (I[2] >= q) ph = 0;
    else if (I[2] <= -q) ph = pi;
    else {
      ph = std::acos(I[2] / q);
      if (I[3] > 0) ph = 6.283185307179586 - ph;
    }
  } else if (I[0] + I[1] <= tol * I[0]) {
    A1 = std::sqrt(I[0]); A2 = 0; ph = 0;
  } else {
    A1 = 0; A2 = std::sqrt(I[0]); ph = 0;
  }
  E[0] = A1; E[1] = A1 * 0.0; E[2] = A2 * std::cos(ph); E[3] = A2 * std::sin(ph);
}

#endif  // MCRT_PART(1)

}  // namespace

// Launchers of the templated kernels, each instantiated in one part (defined
// at the end of this file).
template <typename R, int GEO> hipError_t coop_launch_geo(const CoopPlan& cp, hipStream_t s);
template <typename R> hipError_t batch_set_lds_r(size_t lds_bytes);
template <typename R>
hipError_t batch_launch_r(const BatchChunk* dchunks, int nchunks, unsigned int* dnext,
                          void* slot_mem, int blocks, size_t lds_bytes, hipStream_t s,
                          int64_t* states, const unsigned int* preempt);
template <typename R>
hipError_t eval_launch_r(const CoopPlan& cp, int what, int64_t n, const double* din, double* dout,
                         hipStream_t s);

#if MCRT_PART(1)

// ------------------------------------------------------------- table cache
struct TableCache {
  struct Entry {
    void* F = nullptr;
    void* S = nullptr;
    void* C = nullptr;
    size_t bytes = 0;
    int refs = 0;
    uint64_t last_use = 0;
  };
  std::mutex mu;
  std::map<std::pair<uint64_t, int>, Entry*> map;  // (content hash, element size)
  size_t budget = 0, bytes = 0;
  uint64_t clock = 0, hits = 0, misses = 0;

  ~TableCache() {
    for (auto& kv : map) {
      (void)hipFree(kv.second->F);
      (void)hipFree(kv.second->S);
      (void)hipFree(kv.second->C);
      delete kv.second;
    }
  }
  // Free unreferenced entries, least recently used first, down to the budget.
  void trim() {
    while (bytes > budget) {
      auto victim = map.end();
      for (auto it = map.begin(); it != map.end(); ++it)
        if (it->second->refs == 0 && (victim == map.end() || it->second->last_use < victim->second->last_use))
          victim = it;
      if (victim == map.end()) return;
      Entry* e = victim->second;
      (void)hipFree(e->F);
      (void)hipFree(e->S);
      (void)hipFree(e->C);
      bytes -= e->bytes;
      delete e;
      map.erase(victim);
    }
  }
};

TableCache* table_cache_create(size_t budget_bytes) {
  auto* c = new TableCache();
  c->budget = budget_bytes;
  return c;
}
void table_cache_destroy(TableCache* c) { delete c; }
void table_cache_release(TableCache* c, const std::vector<const void*>& refs) {
  if (!c) return;
  std::lock_guard<std::mutex> g(c->mu);
  for (const void* r : refs) {
    auto* e = (TableCache::Entry*)r;
    if (e->refs > 0) --e->refs;
  }
  c->trim();
}
void table_cache_stats(TableCache* c, size_t* bytes, uint64_t* hits, uint64_t* misses) {
  std::lock_guard<std::mutex> g(c->mu);
  if (bytes) *bytes = c->bytes;
  if (hits) *hits = c->hits;
  if (misses) *misses = c->misses;
}

namespace {

// FNV-1a over the scatterer's sizes and tables: the cache key.
uint64_t table_hash(const mcrt_scatterer* s) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&h](const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
  };
  mix(&s->kind, sizeof s->kind);
  mix(&s->n_theta, sizeof s->n_theta);
  mix(&s->random_points, sizeof s->random_points);
  const size_t N = (size_t)s->n_theta;
  mix(s->F11, 8 * N); mix(s->F12, 8 * N); mix(s->F33, 8 * N); mix(s->F34, 8 * N);
  mix(s->S11, 16 * N); mix(s->S22, 16 * N);
  if (s->kind == MCRT_SCATTERER_SPLINE) { mix(s->F22, 8 * N); mix(s->F44, 8 * N); }
  mix(s->csrn, 8 * ((size_t)s->random_points + 1));
  return h;
}

// Upload a scatterer's tables in R, F and S padded with one zero node: into
// memory of the solve (alloc), or shared through the cache (a reference is
// appended to refs).
template <typename R>
void fill_sca(ScaView<R>& v, const mcrt_scatterer* s, const DevAlloc& alloc, int* rc, std::string* err,
              TableCache* cache = nullptr, std::vector<const void*>* refs = nullptr) {
  v.kind = s->kind == MCRT_SCATTERER_MIE ? kMie : (s->kind == MCRT_SCATTERER_SPLINE ? kSpl : kRay);
  v.Nt = s->n_theta;
  v.rp = s->random_points;
  v.scale = v.kind ? (R)((s->n_theta - 1) / 3.141592653589793) : R(0);
  v.csrn = v.F = v.S = nullptr;
  if (!v.kind) return;
  std::unique_lock<std::mutex> lk;
  std::pair<uint64_t, int> key;
  if (cache) {
    lk = std::unique_lock<std::mutex>(cache->mu);
    key = {table_hash(s), (int)sizeof(R)};
    auto it = cache->map.find(key);
    if (it != cache->map.end()) {
      TableCache::Entry* e = it->second;
      ++e->refs;
      e->last_use = ++cache->clock;
      ++cache->hits;
      refs->push_back(e);
      v.F = (const R*)e->F;
      v.S = (const R*)e->S;
      v.csrn = (const R*)e->C;
      return;
    }
    ++cache->misses;
  }
  const int N = s->n_theta;
  // (F11, F12, F33, F34) per node over N + 1 nodes (the last zero); a Spline's
  // (F22, F44) per node follow at spl_b(N) (ScaView)
  const int fs = f_stride(v.kind);
  std::vector<R> F((size_t)fs * (N + 1), R(0)), S(4 * (size_t)(N + 1), R(0)),
      C((size_t)s->random_points + 1);
  for (int i = 0; i < N; ++i) {
    R* f = &F[4 * (size_t)i];
    f[0] = (R)s->F11[i]; f[1] = (R)s->F12[i]; f[2] = (R)s->F33[i]; f[3] = (R)s->F34[i];
    if (v.kind == kSpl) {
      R* b = &F[(size_t)spl_b(N) + 2 * (size_t)i];
      b[0] = (R)s->F22[i]; b[1] = (R)s->F44[i];
    }
    S[4 * i] = (R)s->S11[2 * i]; S[4 * i + 1] = (R)s->S11[2 * i + 1];
    S[4 * i + 2] = (R)s->S22[2 * i]; S[4 * i + 3] = (R)s->S22[2 * i + 1];
  }
  for (int i = 0; i <= s->random_points; ++i) C[i] = (R)s->csrn[i];
  auto up = [&](const std::vector<R>& h) -> const R* {
    hipError_t e;
    R* d = nullptr;
    if (cache) {
      e = hipMalloc((void**)&d, sizeof(R) * h.size());
      if (e != hipSuccess) d = nullptr;
    } else {
      d = (R*)alloc(sizeof(R) * h.size(), &e);
    }
    if (!d) { *rc = MCRT_ERR_NOMEM; *err = "coop: table allocation failed"; return nullptr; }
    if (hipMemcpy(d, h.data(), sizeof(R) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
      *rc = MCRT_ERR_HIP; *err = "coop: table upload failed"; return nullptr;
    }
    return d;
  };
  v.F = up(F);
  if (v.F) v.S = up(S);
  if (v.S) v.csrn = up(C);
  if (cache) {
    if (!v.csrn) {
      (void)hipFree((void*)v.F);
      (void)hipFree((void*)v.S);
      return;
    }
    auto* e = new TableCache::Entry();
    e->F = (void*)v.F;
    e->S = (void*)v.S;
    e->C = (void*)v.csrn;
    e->bytes = sizeof(R) * (F.size() + S.size() + C.size());
    e->refs = 1;
    e->last_use = ++cache->clock;
    cache->map[key] = e;
    cache->bytes += e->bytes;
    refs->push_back(e);
    cache->trim();
  }
}

// LDS bytes of one workgroup (the layout of coop_kernel).
template <typename R> size_t lds_need(const CoopParams<R>& P) {
  const bool cb = P.compute_cb != 0;
  switch (P.sk) {
    case kRay: return cb ? lds_offsets<R, kRay, true>(P).end : lds_offsets<R, kRay, false>(P).end;
    case kMie: return cb ? lds_offsets<R, kMie, true>(P).end : lds_offsets<R, kMie, false>(P).end;
    default: return cb ? lds_offsets<R, kAny, true>(P).end : lds_offsets<R, kAny, false>(P).end;
  }
}

// The LDS shape of a solve whose table fields are set: the most slots (a
// multiple of 64, at most Shape<R>::kSlots and at most max_slots when > 0) whose
// records fit beside the tables in lds_max bytes, with the S table staged when
// want_s and it fits at that many slots (first choice), else read from HBM.
// Returns false when even 64 slots with S in HBM do not fit.
template <typename R> bool fit_lds(CoopParams<R>& P, size_t lds_max, int max_slots, bool want_s,
                                   size_t* lds_out) {
  const int top = max_slots > 0 ? std::min(Shape<R>::kSlots, std::max(64, max_slots / 64 * 64))
                                : Shape<R>::kSlots;
  const bool can_s = Shape<R>::kStageS && P.compute_cb && P.sca.kind != kRay && want_s;
  for (int stage = can_s ? 1 : 0; stage >= 0; --stage)
    for (int slots = top; slots >= 64; slots -= 64) {
      P.slots = slots;
      P.stage_s = stage;
      const size_t b = lds_need(P);
      // keep the staged S only if it leaves the full slot count
      if (b <= lds_max && (stage == 0 || slots == top)) {
        *lds_out = b;
        return true;
      }
    }
  return false;
}

// Kernel kind of a solve: kRay / kMie when one Rayleigh or Mie scatterer fills
// every layer, else kAny (layers with different scatterers, or a Spline).
inline int solve_kind(const mcrt_geometry* g) {
  const mcrt_scatterer* s0 = g->kind == MCRT_GEOMETRY_PLANE ? g->layers[0].scatterer : g->scatterer;
  const int nl = g->kind == MCRT_GEOMETRY_PLANE ? g->n_layers : 1;
  bool same = true, spline = s0->kind == MCRT_SCATTERER_SPLINE;
  for (int i = 1; i < nl; ++i) {
    same = same && g->layers[i].scatterer == s0;
    spline = spline || g->layers[i].scatterer->kind == MCRT_SCATTERER_SPLINE;
  }
  if (spline || !same) return kAny;
  return s0->kind == MCRT_SCATTERER_MIE ? kMie : kRay;
}

// Lookup constants of a table of n_theta nodes.
template <typename R> TabC<R> tab_consts(const ScaView<R>& v) {
  TabC<R> t{};
  if (v.kind == kRay) return t;
  const double sc = (v.Nt - 1) / 3.141592653589793;
  // acos(a) = √(1-a)·Σ cₖ aᵏ on [0, 1]: the degree-6 minimax fit printed by
  // csrc/fit_acos.py, highest degree first
  static const double c[kAcosTerms] = {0.0026117211109719563, -0.012003396496319869,
                                       0.027762914927838391,  -0.049197437593977882,
                                       0.088835885904433951,  -0.21459108867891987,
                                       1.5707962388530227};
  t.scale = (R)sc;
  for (int k = 0; k < kAcosTerms; ++k) t.acs[k] = (R)(c[k] * sc);
  t.pis = (R)(v.Nt - 1);  // π·sc
  t.boff = spl_b(v.Nt);
  return t;
}

// Fill the parameter block of one solve and upload its tables and bin
// constants (everything but the launch resources: slabs and slot state).
// P.slots is the most slots the solve's LDS allows (at most max_slots when
// > 0), *lds_out the LDS bytes at that count.
template <typename R>
int init_params(CoopParams<R>& P, size_t* lds_out, size_t lds_max, const mcrt_config* cfg,
                const mcrt_geometry* g, const double I0[4], uint64_t seed, uint64_t ray0,
                uint32_t instance, const DevAlloc& alloc, unsigned long long* dcnt,
                std::string* err, TableCache* cache = nullptr,
                std::vector<const void*>* refs = nullptr, int max_slots = 0, bool want_s = true) {
  std::memset(&P, 0, sizeof P);
  // bin offsets are 32-bit in the kernels (kBinStride·bin < 2³¹), and the
  // private FP64 slabs of a larger grid would not fit in device memory
  if ((int64_t)cfg->n_theta * cfg->n_phi > kMaxCoopBins ||
      (cfg->compute_cb && (int64_t)cfg->n_theta_b * cfg->n_phi_b > kMaxCoopBins)) {
    *err = "bin grids of more than 2^26 bins are not supported by the cooperative forms "
           "(use fp64_plain)";
    return MCRT_ERR_UNSUPPORTED;
  }
  const bool plane = g->kind == MCRT_GEOMETRY_PLANE;
  const mcrt_scatterer* s0 = plane ? g->layers[0].scatterer : g->scatterer;
  const mcrt_scatterer* s1 = plane ? g->layers[g->n_layers - 1].scatterer : s0;
  const bool cb = cfg->compute_cb != 0;
  P.sk = solve_kind(g);
  P.compute_cb = cb ? 1 : 0;
  P.lo_shared = s1 == s0;
  {
    // the table fields the LDS layout reads, before anything is uploaded
    auto shape_of = [](ScaView<R>& v, const mcrt_scatterer* s) {
      v.kind = s->kind == MCRT_SCATTERER_MIE ? kMie : (s->kind == MCRT_SCATTERER_SPLINE ? kSpl : kRay);
      v.Nt = s->n_theta;
      v.rp = s->random_points;
    };
    shape_of(P.sca, s0);
    shape_of(P.sca_lo, s1);
    if (!fit_lds(P, lds_max, max_slots, want_s, lds_out)) {
      *err = "the scatterer tables do not fit in LDS beside the event records of 64 ray slots "
             "(reduce Nθ or random_points, or use fp64_plain)";
      return MCRT_ERR_UNSUPPORTED;
    }
  }
  int layer_slot[kMaxLayers] = {};
  GeoDev gd = make_geo(g, layer_slot);
  P.N = cfg->number_of_rays;
  P.maxscat = cfg->maximum_scattering_times;
  P.min_int = (R)cfg->minimum_intensity;
  P.tau0 = (R)cfg->tau0;
  P.compute_cb = cb ? 1 : 0;
  P.geo_kind = gd.kind;
  P.nl = gd.nl;
  for (int i = 0; i < gd.nl; ++i) {
    P.tau[i] = (R)gd.tau[i]; P.omega[i] = (R)gd.omega[i]; P.isinf[i] = gd.isinf[i];
    P.tau_d[i] = gd.tau[i];
  }
  P.kl1 = gd.kl1;
  int rc = MCRT_OK;
  {
    // tables: set 0, set 1 and (kAny) every layer's, each distinct scatterer once
    std::vector<std::pair<const mcrt_scatterer*, ScaView<R>>> done;
    auto view = [&](const mcrt_scatterer* sc, ScaView<R>& v) {
      for (auto& d : done)
        if (d.first == sc) { v = d.second; return; }
      fill_sca<R>(v, sc, alloc, &rc, err, cache, refs);
      if (!rc) done.push_back({sc, v});
    };
    const int slots = P.slots, stage_s = P.stage_s;
    view(s0, P.sca);
    if (!rc) view(s1, P.sca_lo);
    if (!rc && P.sk == kAny)
      for (int i = 0; i < gd.nl && !rc; ++i) view(plane ? g->layers[i].scatterer : s0, P.lsca[i]);
    if (rc) return rc;
    P.slots = slots;
    P.stage_s = stage_s;
  }
  fill_sca<double>(P.sca64, s0, alloc, &rc, err, cache, refs);
  if (rc) return rc;
  {
    P.tc[0] = tab_consts(P.sca);
    P.tc[1] = tab_consts(P.sca_lo);
    P.tau0k = (R)(cfg->tau0 * (double)ExpK<R>::k);
    // inside a sphere every distance to the boundary is at most the diameter
    P.rt_cut = !(g->kind == MCRT_GEOMETRY_SPHERE && 2.0 * g->tau * (1.0 + 1e-6) < cfg->tau0);
    P.cb_cut = !(g->kind == MCRT_GEOMETRY_SPHERE && 4.0 * g->tau * (1.0 + 1e-6) < cfg->tau0);
  }
  const int Nt = (int)cfg->n_theta, Np = (int)cfg->n_phi;
  const int Ntb = cb ? (int)cfg->n_theta_b : 0, Npb = cb ? (int)cfg->n_phi_b : 0;
  P.Nt = Nt; P.Np = Np; P.Ntb = Ntb; P.Npb = Npb;
  if (g->kind == MCRT_GEOMETRY_PLANE) {
    // lower_theta_range 1:Nθ÷2 (transmission), upper Nθ÷2+1:Nθ (reflection)
    P.nranges = 0;
    if (Nt / 2 > 0) { P.range_b[P.nranges] = 0; P.range_e[P.nranges] = Nt / 2; P.range_kind[P.nranges] = 0; ++P.nranges; }
    P.range_b[P.nranges] = Nt / 2; P.range_e[P.nranges] = Nt; P.range_kind[P.nranges] = 1; ++P.nranges;
  } else {
    P.nranges = 1; P.range_b[0] = 0; P.range_e[0] = Nt; P.range_kind[0] = 1;
  }
  // incident frame and CB constants (FP64, host)
  double k0[3], f0[4];
  incident(g, k0, f0);
  for (int i = 0; i < 3; ++i) { P.k0[i] = (R)k0[i]; P.kf0[i] = -k0[i]; }
  for (int i = 0; i < 4; ++i) { P.f0[i] = f0[i]; P.I0[i] = (R)I0[i]; }
  h_E_from_I(I0, P.Ei);
  {
    double a = std::hypot(P.Ei[0], P.Ei[1]), b = std::hypot(P.Ei[2], P.Ei[3]);
    P.norm0 = std::sqrt(a * a + b * b);
  }
  HFr kfs = h_sph(P.kf0);
  // RT bin constants: direction (cpu.jl:13), basis (get_basis), inorm, 1/k_z
  std::vector<R> rtb(kBinStride * (size_t)Nt * Np);
  for (int j2 = 0; j2 < Np; ++j2)
    for (int j1 = 0; j1 < Nt; ++j1) {
      double k[3] = {cfg->sin_theta[j1] * cfg->cos_phi[j2], cfg->sin_theta[j1] * cfg->sin_phi[j2],
                     cfg->cos_theta[j1]};
      double Eh[3], Ev[3];
      h_get_basis(k, Eh, Ev);
      R* b = &rtb[kBinStride * ((size_t)j1 + (size_t)Nt * j2)];
      b[0] = (R)k[0]; b[1] = (R)k[1]; b[2] = (R)k[2]; b[3] = (R)Eh[0]; b[4] = (R)Eh[1];
      b[5] = (R)(Eh[1] * k[2]); b[6] = (R)(-Eh[0] * k[2]); b[7] = (R)(Eh[0] * k[1] - Eh[1] * k[0]);
      b[8] = (R)cfg->inorm[j1];
      b[9] = (R)(1.0 / k[2]);
    }
  // CB bin constants: kf = ray_to_norm(bin, frame(-k₀)), its basis, the final
  // rotation (cpu.jl:118-124), k_i + k_f in FP64
  std::vector<R> cbb(kBinStride * (size_t)std::max(1, Ntb * Npb));
  std::vector<double> kik(4 * (size_t)std::max(1, Ntb * Npb));
  for (int j2 = 0; j2 < Npb; ++j2)
    for (int j1 = 0; j1 < Ntb; ++j1) {
      double v[3] = {cfg->sin_theta_b[j1] * cfg->cos_phi_b[j2],
                     cfg->sin_theta_b[j1] * cfg->sin_phi_b[j2], cfg->cos_theta_b[j1]};
      double kf[3];
      h_ray_to_norm(v, kfs, kf);
      double Eh[3], Ev[3], Ehr[3], Evr[3];
      h_get_basis(kf, Eh, Ev);
      h_norm_to_ray(Eh, kfs, Ehr);
      h_norm_to_ray(Ev, kfs, Evr);
      double cps = cfg->sin_phi_b[j2] * Ehr[0] - cfg->cos_phi_b[j2] * Ehr[1];
      double sps = -cfg->sin_phi_b[j2] * Evr[0] + cfg->cos_phi_b[j2] * Evr[1];
      const size_t lb = (size_t)j1 + (size_t)Ntb * j2;
      R* b = &cbb[kBinStride * lb];
      b[0] = (R)kf[0]; b[1] = (R)kf[1]; b[2] = (R)kf[2]; b[3] = (R)Eh[0]; b[4] = (R)Eh[1];
      b[5] = (R)(Eh[1] * kf[2]); b[6] = (R)(-Eh[0] * kf[2]);
      b[7] = (R)(Eh[0] * kf[1] - Eh[1] * kf[0]);
      b[8] = (R)(cps * cps - sps * sps);
      b[9] = (R)(2.0 * cps * sps);
      b[10] = (R)(1.0 / kf[2]);
      // kl·(k_i + k_f) in the unit of phase_cs (cpu.jl:94)
      const double ku = gd.kl1 * kRevPerRadian;
      kik[4 * lb] = ku * (k0[0] + kf[0]); kik[4 * lb + 1] = ku * (k0[1] + kf[1]);
      kik[4 * lb + 2] = ku * (k0[2] + kf[2]);
      kik[4 * lb + 3] = 0;
    }
  auto up = [&](const void* h, size_t bytes) -> void* {
    hipError_t e;
    void* d = alloc(bytes, &e);
    if (!d) { rc = MCRT_ERR_NOMEM; *err = "coop: allocation failed"; return nullptr; }
    if (hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) {
      rc = MCRT_ERR_HIP; *err = "coop: upload failed"; return nullptr;
    }
    return d;
  };
  P.rt_bin = (const R*)up(rtb.data(), sizeof(R) * rtb.size());
  if (rc) return rc;
  P.cb_bin = (const R*)up(cbb.data(), sizeof(R) * cbb.size());
  if (rc) return rc;
  P.cb_kik = (const double*)up(kik.data(), sizeof(double) * kik.size());
  if (rc) return rc;
  P.seed = seed;
  P.ray0 = ray0;
  P.instance = instance;
  P.nrt = 4 * (int64_t)Nt * Np;
  P.ncb = 4 * (int64_t)cfg->n_theta_b * cfg->n_phi_b;
  P.nout = 5 + P.nrt + P.ncb;
  {
    // the fewest groups g ≤ 8 minimising ⌈g·nb/threads⌉/g (the CB pass time)
    const int nb = Ntb * Npb;
    int best = 1;
    double tbest = 1e300;
    for (int g = 1; g <= 8 && nb > 0; ++g) {
      const double t = (double)((g * nb + kThreads - 1) / kThreads) / g;
      if (t < tbest - 1e-12) { tbest = t; best = g; }
    }
    P.cb_groups = best;
  }
  P.counters = dcnt;
  return MCRT_OK;
}

template <typename R>
int create_impl(CoopPlan& cp, int cu_count, size_t lds_max, const mcrt_config* cfg,
                const mcrt_geometry* g, const double I0[4], uint64_t seed, uint64_t ray0,
                uint32_t instance, const mcrt_launch_shape& shape, const DevAlloc& alloc,
                unsigned long long* dcnt, std::string* err, TableCache* cache) {
  CoopParams<R> P;
  size_t lds = 0;
  int rc = init_params<R>(P, &lds, lds_max, cfg, g, I0, seed, ray0, instance, alloc, dcnt, err,
                          cache, &cp.table_refs, shape.slots, shape.stage_s >= 0);
  if (rc) return rc;
  const bool cb = P.compute_cb != 0;
  auto up = [&](const void* h, size_t bytes) -> void* {
    hipError_t e;
    void* d = alloc(bytes, &e);
    if (!d) { rc = MCRT_ERR_NOMEM; *err = "coop: allocation failed"; return nullptr; }
    if (hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) {
      rc = MCRT_ERR_HIP; *err = "coop: upload failed"; return nullptr;
    }
    return d;
  };
  // launch shape: active slots per workgroup (at most what LDS allows, P.slots
  // from init_params; by default all of them, fewer when the solve has too few
  // rays to give every CU a workgroup), the LDS it needs, and workgroups per
  // CU (LDS bounds how many are resident)
  int slots = P.slots;
  if (shape.slots <= 0 && (P.N + slots - 1) / slots < cu_count) {
    const int64_t per = (P.N + cu_count - 1) / cu_count;
    slots = (int)std::min<int64_t>(slots, std::max<int64_t>(64, (per + 63) / 64 * 64));
  }
  P.slots = slots;
  lds = lds_need(P);
  cp.lds_bytes = lds;
  cp.lds_max = lds_max;
  const int wg_per_cu = std::max<int>(1, (int)(lds_max / std::max<size_t>(lds, 1)));
  const int per_cu = shape.workgroups_per_cu > 0 ? std::min(shape.workgroups_per_cu, 4)
                                                 : std::min(wg_per_cu, 2);
  cp.slots = slots;
  cp.stage_s = P.stage_s;
  cp.wg_per_cu = per_cu;
  const int64_t want = (int64_t)cu_count * per_cu;
  const int64_t byrays = (P.N + slots - 1) / slots;
  cp.blocks = (int)std::max<int64_t>(1, std::min(want, byrays));
  cp.threads = kThreads;
  hipError_t e;
  P.nout = 5 + P.nrt + (int64_t)P.cb_groups * P.ncb;  // private slabs: a CB copy per group
  cp.cb_groups = P.cb_groups;
  cp.slab_words = P.nout;
  cp.nslab = cp.blocks;
  cp.slab_stride = 2 * P.nout;
  size_t slab_bytes = sizeof(double) * (size_t)cp.nslab * (size_t)cp.slab_stride;
  cp.slabs = (double*)alloc(slab_bytes, &e);
  if (!cp.slabs) { *err = "coop: slab allocation failed"; return MCRT_ERR_NOMEM; }
  P.slabs = cp.slabs;
  P.slot_mem = alloc(slot_block_bytes<R>() * (size_t)cp.blocks, &e);
  if (!P.slot_mem) { *err = "coop: slot-state allocation failed"; return MCRT_ERR_NOMEM; }
  cp.nout = 5 + P.nrt + P.ncb;
  cp.nrt = P.nrt;
  cp.ncb = P.ncb;
  cp.params_dev = up(&P, sizeof P);
  if (rc) return rc;
  cp.geo_kind = P.geo_kind;
  cp.sk = P.sk;
  cp.sk0 = P.sca.kind;
  cp.cb = cb;
  return MCRT_OK;
}

}  // namespace

int coop_create(CoopPlan& cp, hipStream_t, int cu_count, size_t lds_max, const mcrt_config* cfg,
                const mcrt_geometry* geom, const double I0[4], uint64_t seed, uint64_t ray0,
                uint32_t instance, bool mixed, const mcrt_launch_shape& shape,
                const DevAlloc& alloc, double* dA, double* dIrt, double* dIcb,
                unsigned long long* dcnt, std::string* err, TableCache* cache) {
  cp.mixed = mixed;
  cp.dA = dA; cp.dIrt = dIrt; cp.dIcb = dIcb;
  if (cu_count <= 0) cu_count = 256;
  if (lds_max == 0) lds_max = 160 * 1024;
  return mixed ? create_impl<float>(cp, cu_count, lds_max, cfg, geom, I0, seed, ray0, instance, shape,
                                    alloc, dcnt, err, cache)
               : create_impl<double>(cp, cu_count, lds_max, cfg, geom, I0, seed, ray0, instance, shape,
                                     alloc, dcnt, err, cache);
}

int table_pin(TableCache* cache, const mcrt_scatterer* s, std::vector<const void*>* refs,
              std::string* err) {
  if (s->kind == MCRT_SCATTERER_RAYLEIGH) return MCRT_OK;
  const DevAlloc none = [](size_t, hipError_t* e) -> void* { *e = hipErrorInvalidValue; return nullptr; };
  int rc = MCRT_OK;
  ScaView<float> vf;
  fill_sca<float>(vf, s, none, &rc, err, cache, refs);
  if (rc) return rc;
  ScaView<double> vd;
  fill_sca<double>(vd, s, none, &rc, err, cache, refs);
  return rc;
}

hipError_t coop_prepare(CoopPlan& cp, hipStream_t s) {
  return hipMemsetAsync(cp.slabs, 0, sizeof(double) * (size_t)cp.nslab * cp.slab_stride, s);
}

hipError_t coop_launch(CoopPlan& cp, hipStream_t s) {
  if (cp.mixed) return cp.geo_kind == 0 ? coop_launch_geo<float, 0>(cp, s) : coop_launch_geo<float, 1>(cp, s);
  return cp.geo_kind == 0 ? coop_launch_geo<double, 0>(cp, s) : coop_launch_geo<double, 1>(cp, s);
}

// ------------------------------------------------------------- batched solves
// Batch slabs per solve: the batch-means σ needs enough batches for bins whose
// per-ray contributions are sparse (a few large values), where few batches
// underestimate the variance.
constexpr int kBatchSlabs = 128;

// Work of one ray in bin evaluations: (RT bins + 3·CB bins + 1) per event,
// with events per ray from a model of the weight: a ray never leaves the
// medium (escaping draws are resampled) and ends when its weight falls below
// minimum_intensity, each event keeping the fraction ω·(1 - p) of it, where
// p, the part that escapes in the peel-off, is taken as 1/(1 + 0.75·τ^1.15)
// (τ the optical thickness of the first layer, or the radius of the sphere;
// fitted to the measured events per ray of the 256-instance batched
// population, τ from 0.1 to 20, which it gives within 0.56-1.67 times for
// 90% of the instances, where 1 + 2τ + τ²/2 gave 0.14-1.34 times); at most
// maximum_scattering_times.
double solve_work_per_ray(const mcrt_config* cfg, const mcrt_geometry* g) {
  const bool plane = g->kind == MCRT_GEOMETRY_PLANE;
  const double tau = plane ? g->layers[0].tau : g->tau;
  const double om = plane ? g->layers[0].omega : g->omega;
  const double maxscat = (double)std::max<int64_t>(1, cfg->maximum_scattering_times);
  const double p = 1.0 / (1.0 + 0.75 * std::pow(std::max(tau, 0.0), 1.15));
  const double keep = std::max(om, 0.0) * (1.0 - p);
  double ev = maxscat;
  if (cfg->minimum_intensity > 0.0 && cfg->minimum_intensity < 1.0 && keep > 0.0 && keep < 1.0)
    ev = std::min(ev, 1.0 + std::log(1.0 / cfg->minimum_intensity) / -std::log(keep));
  else if (keep <= 0.0)
    ev = 1.0;
  const double bins = (double)cfg->n_theta * (double)cfg->n_phi +
                      (cfg->compute_cb ? 3.0 * (double)cfg->n_theta_b * (double)cfg->n_phi_b : 0.0) +
                      1.0;
  return ev * bins;
}

// One workgroup runs a chunk, and the rays of a chunk fill its slots only
// until the chunk's last rays are in flight: the tail of a chunk, where a few
// long rays keep a workgroup iterating with most slots empty, costs ~25-40 µs
// per iteration whatever the chunk size.
//  * Throughput (batches): ⌈N/64⌉ rays, 512 to 8192, so that tails are a small
//    part of a chunk and a solve still spreads over many workgroups.
//  * Latency (the server): a solve ends no sooner than its longest chunk, so
//    a chunk holds about kChunkWork bin evaluations (some 20 ms of one CU in
//    the mixed form), 64 to 8192 rays.
// Both give at least two chunks when N ≥ 2 (a standard error).
constexpr double kChunkWork = 4e7;

int64_t default_chunk_rays(int64_t N, double work_per_ray, ChunkRule rule) {
  int64_t c;
  if (rule == ChunkRule::Latency) {
    const double w = std::ceil(kChunkWork / std::max(work_per_ray, 1.0));
    c = (int64_t)std::max(64.0, std::min(8192.0, w));
  } else {
    c = std::max<int64_t>(512, std::min<int64_t>(8192, (N + 63) / 64));
  }
  if (N >= 2 && c >= N) c = (N + 1) / 2;
  return std::max<int64_t>(1, c);
}

namespace {
template <typename R>
int solve_create_impl(SolveDev& sd, size_t lds_max, const mcrt_config* cfg, const mcrt_geometry* g,
                      const double I0[4], uint64_t seed, uint64_t ray0, uint32_t instance,
                      int64_t chunk_rays, ChunkRule rule, const DevAlloc& alloc, std::string* err,
                      TableCache* cache, int slots, int stage_s) {
  hipError_t e;
  sd.counters = (unsigned long long*)alloc(sizeof(unsigned long long) * kCntN, &e);
  if (!sd.counters) { *err = "batch: counter allocation failed"; return MCRT_ERR_NOMEM; }
  CoopParams<R> P;
  size_t lds = 0;
  int rc = init_params<R>(P, &lds, lds_max, cfg, g, I0, seed, ray0, instance, alloc, sd.counters, err,
                          cache, &sd.table_refs, slots, stage_s >= 0);
  if (rc) return rc;
  // P.slots: the most slots the solve's LDS allows
  sd.mixed = sizeof(R) == 4;
  sd.geo_kind = P.geo_kind;
  sd.cb = P.compute_cb != 0;
  sd.N = P.N;
  sd.nout = P.nout;
  sd.nrt = P.nrt;
  sd.ncb = P.ncb;
  sd.lds_bytes = lds;
  const double work = solve_work_per_ray(cfg, g);
  sd.chunk_rays = chunk_rays > 0 ? chunk_rays : default_chunk_rays(P.N, work, rule);
  sd.nchunks = (P.N + sd.chunk_rays - 1) / sd.chunk_rays;
  sd.nbatch = (int)std::min<int64_t>(kBatchSlabs, sd.nchunks);
  std::vector<int64_t> br(sd.nbatch, 0);
  for (int64_t k = 0; k < sd.nchunks; ++k)
    br[k % sd.nbatch] += std::min(P.N, (k + 1) * sd.chunk_rays) - k * sd.chunk_rays;
  sd.slabs = (double*)alloc(sizeof(double) * (size_t)sd.nbatch * (size_t)sd.nout, &e);
  sd.batch_rays = (int64_t*)alloc(sizeof(int64_t) * br.size(), &e);
  sd.out = (double*)alloc(sizeof(double) * (size_t)sd.nout, &e);
  sd.sig = (double*)alloc(sizeof(double) * (size_t)sd.nout, &e);
  void* pd = alloc(sizeof P, &e);
  if (!sd.slabs || !sd.batch_rays || !sd.out || !sd.sig || !pd) {
    *err = "batch: solve allocation failed";
    return MCRT_ERR_NOMEM;
  }
  if (hipMemcpy(sd.batch_rays, br.data(), sizeof(int64_t) * br.size(), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(pd, &P, sizeof P, hipMemcpyHostToDevice) != hipSuccess) {
    *err = "batch: upload failed";
    return MCRT_ERR_HIP;
  }
  sd.params_dev = pd;
  // relative work of one chunk (only the order of chunks depends on it)
  sd.cost = (double)sd.chunk_rays * work;
  return MCRT_OK;
}
}  // namespace

int solve_create(SolveDev& sd, size_t lds_max, const mcrt_config* cfg, const mcrt_geometry* geom,
                 const double I0[4], uint64_t seed, uint64_t ray0, uint32_t instance, bool mixed,
                 int64_t chunk_rays, ChunkRule rule, const DevAlloc& alloc, std::string* err,
                 TableCache* cache, int slots, int stage_s) {
  if (lds_max == 0) lds_max = 160 * 1024;
  lds_max -= 16;  // the batch kernel's static LDS (the chunk index)
  return mixed ? solve_create_impl<float>(sd, lds_max, cfg, geom, I0, seed, ray0, instance,
                                          chunk_rays, rule, alloc, err, cache, slots, stage_s)
               : solve_create_impl<double>(sd, lds_max, cfg, geom, I0, seed, ray0, instance,
                                           chunk_rays, rule, alloc, err, cache, slots, stage_s);
}

size_t solve_bytes_estimate(const mcrt_config* cfg, const mcrt_geometry* g, bool mixed,
                            int64_t chunk_rays, ChunkRule rule) {
  const int64_t N = cfg->number_of_rays;
  const int64_t cr =
      chunk_rays > 0 ? chunk_rays : default_chunk_rays(N, solve_work_per_ray(cfg, g), rule);
  const int64_t nb = std::min<int64_t>(kBatchSlabs, (N + cr - 1) / cr);
  const int64_t nbins = (int64_t)cfg->n_theta * cfg->n_phi +
                        (cfg->compute_cb ? (int64_t)cfg->n_theta_b * cfg->n_phi_b : 0);
  const int64_t nout = 5 + 4 * nbins;
  const size_t r = mixed ? 4 : 8;
  return (size_t)(nb + 2) * (size_t)nout * 8 + (size_t)nbins * kBinStride * r +
         (size_t)nbins * 32 + sizeof(CoopParams<double>) + 4096;
}

void solve_chunk(const SolveDev& sd, int64_t k, BatchChunk* out) {
  out->params = sd.params_dev;
  out->rb = k * sd.chunk_rays;
  out->re = std::min(sd.N, (k + 1) * sd.chunk_rays);
  out->slab = sd.slabs + (size_t)(k % sd.nbatch) * (size_t)sd.nout;
  out->cont = nullptr;
  out->resume = 0;
  out->pad = 0;
}

static hipError_t zero(void* p, size_t bytes, hipStream_t s) {
  const int64_t n = (int64_t)(bytes / 4);
  const int bl = (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + 255) / 256));
  hipLaunchKernelGGL(zero_words, dim3(bl), dim3(256), 0, s, (uint32_t*)p, n);
  return hipGetLastError();
}

hipError_t solve_prepare(const SolveDev& sd, hipStream_t s) {
  hipError_t e = zero(sd.slabs, sizeof(double) * (size_t)sd.nbatch * (size_t)sd.nout, s);
  if (e == hipSuccess) e = zero(sd.counters, sizeof(unsigned long long) * kCntN, s);
  return e;
}

hipError_t solve_finish(const SolveDev& sd, hipStream_t s) {
  const int th = 256;
  const int bl = (int)((sd.nout + th - 1) / th);
  hipLaunchKernelGGL(batch_reduce, dim3(bl), dim3(th), 0, s, (const double*)sd.slabs, sd.nbatch,
                     sd.nout, sd.N, (const int64_t*)sd.batch_rays, sd.out, sd.sig);
  return hipGetLastError();
}

size_t batch_slot_bytes(bool mixed, int blocks) {
  return mixed ? slot_block_bytes<float>() * (size_t)blocks
               : slot_block_bytes<double>() * (size_t)blocks;
}

hipError_t batch_set_lds(bool mixed, size_t lds_bytes) {
  return mixed ? batch_set_lds_r<float>(lds_bytes) : batch_set_lds_r<double>(lds_bytes);
}

hipError_t batch_launch(bool mixed, const BatchChunk* dchunks, int nchunks, unsigned int* dnext,
                        void* slot_mem, int blocks, size_t lds_bytes, hipStream_t s, int64_t* states,
                        const unsigned int* preempt, bool zero_next) {
  if (zero_next) {
    hipError_t e = zero(dnext, sizeof(unsigned int), s);
    if (e != hipSuccess) return e;
  }
  return mixed ? batch_launch_r<float>(dchunks, nchunks, dnext, slot_mem, blocks, lds_bytes, s,
                                       states, preempt)
               : batch_launch_r<double>(dchunks, nchunks, dnext, slot_mem, blocks, lds_bytes, s,
                                        states, preempt);
}

hipError_t coop_sigma(CoopPlan& cp, hipStream_t s, int64_t N, double* sig) {
  const int th = 256;
  const int bl = (int)((cp.nout + th - 1) / th);
  hipLaunchKernelGGL(slab_sigma, dim3(bl), dim3(th), 0, s, (const double*)cp.slabs, cp.nslab,
                     cp.slab_words, cp.nout, N, cp.nrt, cp.ncb, cp.cb_groups, sig);
  return hipGetLastError();
}

int coop_eval(const CoopPlan& cp, hipStream_t s, int what, int64_t n, const double* in,
              double* out, std::string* err) {
  if (n <= 0) return MCRT_OK;
  std::vector<double> h(in, in + 16 * (size_t)n);
  if (what == MCRT_EVAL_UPDATE_I || what == MCRT_EVAL_UPDATE_E) {
    // the bin constants of k, as init_params tabulates them
    for (int64_t i = 0; i < n; ++i) {
      double* a = &h[16 * (size_t)i];
      double Eh[3], Ev[3];
      h_get_basis(a + 8, Eh, Ev);
      a[11] = Eh[0]; a[12] = Eh[1];
      a[13] = Eh[1] * a[10]; a[14] = -Eh[0] * a[10]; a[15] = Eh[0] * a[9] - Eh[1] * a[8];
    }
  }
  void *din = nullptr, *dout = nullptr;
  hipError_t e = hipMalloc(&din, sizeof(double) * h.size());
  if (e == hipSuccess) e = hipMalloc(&dout, sizeof(double) * 4 * (size_t)n);
  if (e == hipSuccess) e = hipMemcpyAsync(din, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = cp.mixed ? eval_launch_r<float>(cp, what, n, (const double*)din, (double*)dout, s)
                 : eval_launch_r<double>(cp, what, n, (const double*)din, (double*)dout, s);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(double) * 4 * (size_t)n, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(din);
  (void)hipFree(dout);
  if (e != hipSuccess) {
    *err = hipGetErrorString(e);
    return MCRT_ERR_HIP;
  }
  return MCRT_OK;
}

hipError_t coop_reduce(CoopPlan& cp, hipStream_t s) {
  const int th = 256;
  const int bl = (int)((cp.nout + th - 1) / th);
  hipLaunchKernelGGL(slab_reduce, dim3(bl), dim3(th), 0, s, (const double*)cp.slabs, cp.nslab,
                     cp.slab_words, cp.nout, cp.dA, cp.dIrt, cp.dIcb, cp.nrt, cp.ncb, cp.cb_groups);
  return hipGetLastError();
}

#endif  // MCRT_PART(1)

// ------------------------------------------------------- kernel launchers
#if MCRT_COOP_PART != 1
namespace {
template <typename R, int GEO, int SK, bool CB>
hipError_t launch_t(const CoopPlan& cp, hipStream_t s) {
  auto k = coop_kernel<R, GEO, SK, CB>;
  // the attribute is per kernel and process-wide: the device maximum, so that
  // plans of other contexts and threads never see a smaller value set here
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)cp.lds_max);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(cp.blocks), dim3(cp.threads), cp.lds_bytes, s,
                     (const CoopParams<R>*)cp.params_dev);
  return hipGetLastError();
}

template <typename R, int GEO, int SK>
hipError_t launch_cb(const CoopPlan& cp, hipStream_t s) {
  return cp.cb ? launch_t<R, GEO, SK, true>(cp, s) : launch_t<R, GEO, SK, false>(cp, s);
}
template <typename R, int GEO>
hipError_t launch_geo(const CoopPlan& cp, hipStream_t s) {
  if (cp.sk == kMie) return launch_cb<R, GEO, kMie>(cp, s);
  if (cp.sk == kRay) return launch_cb<R, GEO, kRay>(cp, s);
  return launch_cb<R, GEO, kAny>(cp, s);
}

}  // namespace

template <typename R, int GEO> hipError_t coop_launch_geo(const CoopPlan& cp, hipStream_t s) {
  return launch_geo<R, GEO>(cp, s);
}

template <typename R> hipError_t batch_set_lds_r(size_t lds_bytes) {
  return hipFuncSetAttribute((const void*)batch_kernel<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds_bytes);
}

template <typename R>
hipError_t batch_launch_r(const BatchChunk* dchunks, int nchunks, unsigned int* dnext,
                          void* slot_mem, int blocks, size_t lds_bytes, hipStream_t s,
                          int64_t* states, const unsigned int* preempt) {
  hipLaunchKernelGGL(batch_kernel<R>, dim3(blocks), dim3(kThreads), lds_bytes, s, dchunks, nchunks,
                     dnext, slot_mem, states, preempt);
  return hipGetLastError();
}

// The building blocks of table set 0 (layer 1's scatterer).
template <typename R>
hipError_t eval_launch_r(const CoopPlan& cp, int what, int64_t n, const double* din, double* dout,
                         hipStream_t s) {
  const int th = 256;
  const int bl = (int)((n + th - 1) / th);
  const auto* P = (const CoopParams<R>*)cp.params_dev;
  if (cp.sk0 == kMie) hipLaunchKernelGGL((eval_kernel<R, kMie>), dim3(bl), dim3(th), 0, s, P, what, n, din, dout);
  else if (cp.sk0 == kSpl) hipLaunchKernelGGL((eval_kernel<R, kSpl>), dim3(bl), dim3(th), 0, s, P, what, n, din, dout);
  else hipLaunchKernelGGL((eval_kernel<R, kRay>), dim3(bl), dim3(th), 0, s, P, what, n, din, dout);
  return hipGetLastError();
}
#endif

#if MCRT_COOP_PART == 2
template hipError_t coop_launch_geo<float, 0>(const CoopPlan&, hipStream_t);
#endif
#if MCRT_COOP_PART == 3
template hipError_t coop_launch_geo<float, 1>(const CoopPlan&, hipStream_t);
#endif
#if MCRT_COOP_PART == 4
template hipError_t batch_set_lds_r<float>(size_t);
template hipError_t batch_launch_r<float>(const BatchChunk*, int, unsigned int*, void*, int, size_t,
                                          hipStream_t, int64_t*, const unsigned int*);
template hipError_t eval_launch_r<float>(const CoopPlan&, int, int64_t, const double*, double*,
                                         hipStream_t);
#endif
#if MCRT_COOP_PART == 5
template hipError_t coop_launch_geo<double, 0>(const CoopPlan&, hipStream_t);
template hipError_t coop_launch_geo<double, 1>(const CoopPlan&, hipStream_t);
template hipError_t batch_set_lds_r<double>(size_t);
template hipError_t batch_launch_r<double>(const BatchChunk*, int, unsigned int*, void*, int, size_t,
                                           hipStream_t, int64_t*, const unsigned int*);
template hipError_t eval_launch_r<double>(const CoopPlan&, int, int64_t, const double*, double*,
                                          hipStream_t);
#endif

}  // namespace mcrt
