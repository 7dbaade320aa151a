// mcrt_device.hpp — device-side statement of the per-ray physics, templated
// over the arithmetic type R (double for the FP64 forms, float for the
// transport and bin evaluation of the mixed form).
//
// Reference functions restated here (src/utils.jl unless noted):
//   spherical_from_cartesian 9-23, ray_to_norm 30-34, norm_to_ray 41-45,
//   get_basis 52-62, rotate 82-84, update_I 91-112, E_from_I 119-150,
//   I_from_E 157-163, update_E 170-192, kepler_solver 243-261,
//   scatter 294-323; Rayleigh phase/amplitude (scatterers/rayleigh.jl:8-23);
//   Mie phase/amplitude lookup (scatterers/mie.jl:46-92); Spline phase/
//   amplitude lookup (scatterers/spline.jl:26-76, NSPL = Nθ - 1).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcrt {

#define MCRT_DEV __device__ __forceinline__
#define MCRT_HD __host__ __device__ __forceinline__

// ---------------------------------------------------------------- math shims
template <typename R> struct M;
template <> struct M<double> {
  static MCRT_DEV double sqrt(double x) { return ::sqrt(x); }
  static MCRT_DEV double rsqrt(double x) { return 1.0 / ::sqrt(x); }
  static MCRT_DEV double exp(double x) { return ::exp(x); }
  static MCRT_DEV double log(double x) { return ::log(x); }
  static MCRT_DEV double acos(double x) { return ::acos(x); }
  static MCRT_DEV double cbrt(double x) { return ::cbrt(x); }
  static MCRT_DEV double sin(double x) { return ::sin(x); }
  static MCRT_DEV double cos(double x) { return ::cos(x); }
  static MCRT_DEV void sincos(double x, double* s, double* c) { ::sincos(x, s, c); }
  static MCRT_DEV double floor(double x) { return ::floor(x); }
  static MCRT_DEV double trunc(double x) { return ::trunc(x); }
  static MCRT_DEV double fabs(double x) { return ::fabs(x); }
  static MCRT_DEV double div(double a, double b) { return a / b; }
  static constexpr double eps = 2.220446049250313e-16;
  static constexpr double pi = 3.141592653589793;
  static constexpr double two_pi = 6.283185307179586;
  static constexpr double four_pi = 12.566370614359172;
  static constexpr double kepler_tol = 1e-12;
};
template <> struct M<float> {
  static MCRT_DEV float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }  // v_sqrt_f32
  static MCRT_DEV float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
  // 2^(x·log2 e) on the hardware exp2 (v_exp_f32).
  static MCRT_DEV float exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
  static MCRT_DEV float log(float x) { return ::logf(x); }
  static MCRT_DEV float acos(float x) { return ::acosf(x); }
  static MCRT_DEV float cbrt(float x) { return ::cbrtf(x); }
  static MCRT_DEV float sin(float x) { return ::sinf(x); }
  static MCRT_DEV float cos(float x) { return ::cosf(x); }
  static MCRT_DEV void sincos(float x, float* s, float* c) { ::sincosf(x, s, c); }
  static MCRT_DEV float floor(float x) { return ::floorf(x); }
  static MCRT_DEV float trunc(float x) { return ::truncf(x); }
  static MCRT_DEV float fabs(float x) { return ::fabsf(x); }
  static MCRT_DEV float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  static constexpr float eps = 1.1920929e-7f;  // FP32 machine epsilon (the reference's eps() is FP64's)
  static constexpr float pi = 3.14159265358979f;
  static constexpr float two_pi = 6.28318530717959f;
  static constexpr float four_pi = 12.5663706143592f;
  static constexpr float kepler_tol = 2e-6f;
};

template <typename R> MCRT_DEV R jl_sign(R x) {
  return x > R(0) ? R(1) : (x < R(0) ? R(-1) : x);
}
template <typename R> MCRT_DEV R clamp1(R x) { return x > R(1) ? R(1) : (x < R(-1) ? R(-1) : x); }
template <typename R> MCRT_DEV R sqrt0(R x) { return M<R>::sqrt(x < R(0) ? R(0) : x); }

// ------------------------------------------------------------------ vectors
template <typename R> struct V3 { R x, y, z; };
template <typename R> struct Fr { R ct, st, cp, sp; };  // (cosθ, sinθ, cosϕ, sinϕ)
template <typename R> struct Cx { R re, im; };

template <typename R> MCRT_DEV V3<R> v3(R a, R b, R c) { return V3<R>{a, b, c}; }
template <typename R> MCRT_DEV V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> MCRT_DEV V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> MCRT_DEV V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <typename R> MCRT_DEV R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

template <typename R> MCRT_DEV Cx<R> cx(R a, R b) { return Cx<R>{a, b}; }
template <typename R> MCRT_DEV Cx<R> operator+(Cx<R> a, Cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> MCRT_DEV Cx<R> operator-(Cx<R> a) { return {-a.re, -a.im}; }
template <typename R> MCRT_DEV Cx<R> operator*(Cx<R> a, Cx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> MCRT_DEV Cx<R> operator*(Cx<R> a, R s) { return {a.re * s, a.im * s}; }
template <typename R> MCRT_DEV R abs2(Cx<R> a) { return a.re * a.re + a.im * a.im; }

// ------------------------------------------------------------------- frames
// spherical_from_cartesian (utils.jl:9-23); |x| = 1 is assumed.
template <typename R> MCRT_DEV Fr<R> sph(V3<R> x) {
  R s2 = x.x * x.x + x.y * x.y;
  Fr<R> f;
  if (s2 < R(1e-24)) {
    f.ct = jl_sign(x.z); f.st = R(0); f.cp = R(1); f.sp = R(0);
  } else {
    f.ct = x.z;
    f.st = sqrt0(R(1) - x.z * x.z);
    R sinv = M<R>::rsqrt(s2);
    f.cp = x.x * sinv;
    f.sp = x.y * sinv;
  }
  return f;
}

// ray_to_norm (utils.jl:30-34)
template <typename R> MCRT_DEV V3<R> ray_to_norm(V3<R> x, Fr<R> f) {
  return {x.x * f.ct * f.cp - x.y * f.sp + x.z * f.st * f.cp,
          x.x * f.ct * f.sp + x.y * f.cp + x.z * f.st * f.sp,
          -x.x * f.st + x.z * f.ct};
}
// norm_to_ray (utils.jl:41-45)
template <typename R> MCRT_DEV V3<R> norm_to_ray(V3<R> x, Fr<R> f) {
  return {x.x * f.ct * f.cp + x.y * f.ct * f.sp - x.z * f.st,
          -x.x * f.sp + x.y * f.cp,
          x.x * f.st * f.cp + x.y * f.st * f.sp + x.z * f.ct};
}

// get_basis (utils.jl:52-62), including Ev₂ = -k₃·Eₕ₁ as written there.
template <typename R> MCRT_DEV void get_basis(V3<R> k, V3<R>& Eh, V3<R>& Ev) {
  if (M<R>::fabs(k.z) > R(1) - M<R>::eps) {
    Eh = {R(0), R(-1), R(0)};
    Ev = {jl_sign(k.z), R(0), R(0)};
  } else {
    R ev3 = -sqrt0(R(1) - k.z * k.z);
    R inv = M<R>::div(R(1), ev3);
    R eh1 = k.y * -inv;
    R eh2 = k.x * inv;
    Eh = {eh1, eh2, R(0)};
    Ev = {-k.z * eh2, -k.z * eh1, ev3};
  }
}

// rotate(I, cos2ϕ, sin2ϕ) (utils.jl:82-84)
template <typename R> MCRT_DEV void rotate(R I[4], R c2, R s2) {
  R a = c2 * I[1] + s2 * I[2];
  R b = -s2 * I[1] + c2 * I[2];
  I[1] = a; I[2] = b;
}

// ------------------------------------------------------------ scatterers
// Device view of a scatterer's tables.  F holds (F11, F12, F33, F34) per node
// (Mie); for a Spline, the same block over Nt + 1 nodes (the last one zero)
// followed by (F22, F44) per node over Nt + 1 nodes (at spl_b(Nt)), so that a
// lookup reads 16 + 8 bytes per node with the Mie table's bank spread.  S holds
// (S11 re, S11 im, S22 re, S22 im) per node; all interleaved.
template <typename R> struct ScaView {
  int kind;    // 0 Rayleigh, 1 Mie, 2 Spline
  int Nt;      // table length
  int rp;      // random_points
  R scale;     // (Nt - 1)/π
  const R* csrn;
  const R* F;
  const R* S;
};
// Words per node of a tabulated scatterer's F table, and the offset of a
// Spline table's (F22, F44) block.
MCRT_HD int f_stride(int kind) { return kind == 2 ? 6 : 4; }
MCRT_HD int spl_b(int Nt) { return 4 * (Nt + 1); }

// Inverse-CDF sample of cosθ from the uniform u (rayleigh.jl:9-11, mie.jl:47-51,
// spline.jl:27-31: the same csrn lookup for Mie and Spline).
// The lookups below keep a*b+c unfused: a fused `(i0 + 1) - u·rp` or
// `(i0 + 1) - θ/π·(Nθ-1)` would pair an unrounded coordinate in one weight with
// the rounded one in the other, so the two weights would not sum to 1 (by up to
// half an ulp of the coordinate: 3e-5 at rn ≈ 1000 and 6e-5 at sn ≈ 1500 in
// FP32).  Unfused, the weights round as the reference evaluates them.
template <typename R> MCRT_DEV R phase_sample(const ScaView<R>& s, R u) {
#pragma clang fp contract(off)
  if (s.kind == 0) {
    R rn = R(2) * (R(1) - R(2) * u);
    R q = M<R>::sqrt(R(1) + rn * rn);
    return M<R>::cbrt(q + rn) - M<R>::cbrt(q - rn);
  }
  R rn = u * R(s.rp);
  int mrn = (int)M<R>::trunc(rn);
  if (mrn > s.rp - 1) mrn = s.rp - 1;
  return (R(mrn + 1) - rn) * s.csrn[mrn] + (rn - R(mrn)) * s.csrn[mrn + 1];
}

// Phase matrix at cosθ = mu: P = (P11, P12, P22, P33, P34, P44) (rayleigh.jl:12-17,
// mie.jl:53-71 and spline.jl:33-57; P22 = P11 and P44 = P33 for Rayleigh and
// Mie).  The table lookups clamp |mu| ≤ 1 before acos.  Mie interpolates
// linearly with weights (msn - sn, sn - msn + 1); Spline keeps its weights
// (msn + 1 - sn, sn - msn) as written, one node step further on.
template <typename R> MCRT_DEV void phase_eval(const ScaView<R>& s, R mu, R P[6]) {
#pragma clang fp contract(off)
  if (s.kind == 0) {
    R m2 = mu * mu;
    P[0] = R(0.75) * (R(1) + m2);
    P[1] = R(-0.75) * (R(1) - m2);
    P[3] = R(1.5) * mu;
    P[4] = R(0);
    P[2] = P[0];
    P[5] = P[3];
    return;
  }
  R th = M<R>::acos(clamp1(mu));
  R sn = th / M<R>::pi * R(s.Nt - 1);
  int i0 = (int)M<R>::trunc(sn);
  i0 = i0 < 0 ? 0 : (i0 > s.Nt - 1 ? s.Nt - 1 : i0);
  if (s.kind == 1) {
    R w0 = R(i0 + 1) - sn;
    const R* a = s.F + 4 * i0;
    P[0] = w0 * a[0]; P[1] = w0 * a[1]; P[3] = w0 * a[2]; P[4] = w0 * a[3];
    if (i0 + 1 <= s.Nt - 1) {
      R w1 = sn - R(i0 + 1) + R(1);
      P[0] += w1 * a[4]; P[1] += w1 * a[5]; P[3] += w1 * a[6]; P[4] += w1 * a[7];
    }
    P[2] = P[0];
    P[5] = P[3];
    return;
  }
  R w0 = R(i0 + 2) - sn;
  const R* a = s.F + 4 * i0;
  const R* b = s.F + spl_b(s.Nt) + 2 * i0;
  P[0] = w0 * a[0]; P[1] = w0 * a[1]; P[2] = w0 * b[0];
  P[3] = w0 * a[2]; P[4] = w0 * a[3]; P[5] = w0 * b[1];
  if (i0 + 1 <= s.Nt - 1) {
    R w1 = sn - R(i0 + 1);
    P[0] += w1 * a[4]; P[1] += w1 * a[5]; P[2] += w1 * b[2];
    P[3] += w1 * a[6]; P[4] += w1 * a[7]; P[5] += w1 * b[3];
  }
}

// Amplitude matrix diag(S11, S22) at cosθ = mu (rayleigh.jl:20-23,
// mie.jl:76-92 and spline.jl:60-76, the same lookup with its one-step-shifted
// weights (msn+1-sn, sn-msn)).
template <typename R> MCRT_DEV void amplitude(const ScaView<R>& s, R mu, Cx<R>& S11, Cx<R>& S22) {
#pragma clang fp contract(off)
  if (s.kind == 0) {
    S11 = {R(0), R(1)};
    S22 = {mu * R(0), mu};
    return;
  }
  R th = M<R>::acos(clamp1(mu));
  R sn = th / M<R>::pi * R(s.Nt - 1);
  int i0 = (int)M<R>::trunc(sn);
  i0 = i0 < 0 ? 0 : (i0 > s.Nt - 1 ? s.Nt - 1 : i0);
  R w0 = R(i0 + 2) - sn;
  const R* a = s.S + 4 * i0;
  S11 = {w0 * a[0], w0 * a[1]};
  S22 = {w0 * a[2], w0 * a[3]};
  if (i0 + 1 <= s.Nt - 1) {
    R w1 = sn - R(i0 + 1);
    S11.re += w1 * a[4]; S11.im += w1 * a[5];
    S22.re += w1 * a[6]; S22.im += w1 * a[7];
  }
}

// ---------------------------------------------------- Stokes / field updates
// update_I (utils.jl:91-112): scatter I from the frame f into direction k.
template <typename R>
MCRT_DEV void update_I(const ScaView<R>& s, const R I[4], V3<R> k, Fr<R> f, R out[4],
                       V3<R>* Eh_out = nullptr, V3<R>* Ev_out = nullptr) {
  V3<R> Eh, Ev;
  get_basis(k, Eh, Ev);
  V3<R> kr = norm_to_ray(k, f);
  Fr<R> ks = sph(kr);
  V3<R> Ehk = ray_to_norm(V3<R>{ks.sp, -ks.cp, R(0)}, f);
  V3<R> Evk = ray_to_norm(V3<R>{ks.ct * ks.cp, ks.ct * ks.sp, -ks.st}, f);
  R cpsi = Eh.x * Ehk.x + Eh.y * Ehk.y;
  R spsi = -Eh.x * Evk.x - Eh.y * Evk.y;
  R P[6];
  phase_eval(s, kr.z, P);
  R Ir[4] = {I[0], I[1], I[2], I[3]};
  rotate(Ir, ks.cp * ks.cp - ks.sp * ks.sp, R(-2) * ks.sp * ks.cp);
  out[0] = P[0] * Ir[0] + P[1] * Ir[1];
  out[1] = P[1] * Ir[0] + P[2] * Ir[1];
  out[2] = P[3] * Ir[2] + P[4] * Ir[3];
  out[3] = -P[4] * Ir[2] + P[5] * Ir[3];
  rotate(out, cpsi * cpsi - spsi * spsi, R(2) * spsi * cpsi);
  if (Eh_out) *Eh_out = Eh;
  if (Ev_out) *Ev_out = Ev;
}

// E_from_I (utils.jl:119-150)
template <typename R> MCRT_DEV void E_from_I(const R I[4], Cx<R>& E1, Cx<R>& E2) {
  const R tol = R(1e-14);
  R A1, A2, phi2;
  if (I[0] - M<R>::fabs(I[1]) > tol * I[0]) {
    A1 = M<R>::sqrt(R(0.5) * (I[0] - I[1]));
    A2 = M<R>::sqrt(R(0.5) * (I[0] + I[1]));
    R q = M<R>::sqrt(I[0] * I[0] - I[1] * I[1]);
    if (I[2] >= q) phi2 = R(0);
    else if (I[2] <= -q) phi2 = M<R>::pi;
    else {
      phi2 = M<R>::acos(I[2] / q);
      if (I[3] > R(0)) phi2 = M<R>::two_pi - phi2;
    }
  } else if (I[0] + I[1] <= tol * I[0]) {
    A1 = M<R>::sqrt(I[0]); A2 = R(0); phi2 = R(0);
  } else {
    A1 = R(0); A2 = M<R>::sqrt(I[0]); phi2 = R(0);
  }
  E1 = {A1, A1 * R(0)};
  E2 = {A2 * M<R>::cos(phi2), A2 * M<R>::sin(phi2)};
}

// I_from_E (utils.jl:157-163)
template <typename R> MCRT_DEV void I_from_E(Cx<R> E1, Cx<R> E2, R o[4]) {
  R qh = E1.re * E1.re + E1.im * E1.im;
  R qv = E2.re * E2.re + E2.im * E2.im;
  R r = E2.re * E1.re + E2.im * E1.im;  // Re(E₂ conj(E₁))
  R i = E2.im * E1.re - E2.re * E1.im;  // Im(E₂ conj(E₁))
  o[0] = qh + qv;
  o[1] = -qh + qv;
  o[2] = r + r;
  o[3] = -(i + i);
}

// The 2×2 complex linear map of update_E (utils.jl:170-192) for scattering
// into direction k from the frame f:  E' = Rψ · diag(S11, S22) · Rϕ · E.
template <typename R> struct Lmap { Cx<R> a, b, c, d; };  // [[a, b], [c, d]]

template <typename R>
MCRT_DEV void update_E(const ScaView<R>& s, Cx<R>& E1, Cx<R>& E2, V3<R> k, Fr<R> f,
                       V3<R>* Eh_out = nullptr, V3<R>* Ev_out = nullptr) {
  V3<R> Eh, Ev;
  get_basis(k, Eh, Ev);
  V3<R> kr = norm_to_ray(k, f);
  Fr<R> ks = sph(kr);
  V3<R> Ehk = ray_to_norm(V3<R>{ks.sp, -ks.cp, R(0)}, f);
  V3<R> Evk = ray_to_norm(V3<R>{ks.ct * ks.cp, ks.ct * ks.sp, -ks.st}, f);
  R cpsi = Eh.x * Ehk.x + Eh.y * Ehk.y;
  R spsi = Eh.x * Evk.x + Eh.y * Evk.y;
  Cx<R> S11, S22;
  amplitude(s, kr.z, S11, S22);
  Cx<R> e1 = E1 * ks.cp + E2 * ks.sp;
  Cx<R> e2 = (-E1) * ks.sp + E2 * ks.cp;
  e1 = S11 * e1;
  e2 = S22 * e2;
  E1 = e1 * cpsi + e2 * spsi;
  E2 = (-e1) * spsi + e2 * cpsi;
  if (Eh_out) *Eh_out = Eh;
  if (Ev_out) *Ev_out = Ev;
}

template <typename R>
MCRT_DEV Lmap<R> update_E_map(const ScaView<R>& s, V3<R> k, Fr<R> f) {
  V3<R> Eh, Ev;
  get_basis(k, Eh, Ev);
  V3<R> kr = norm_to_ray(k, f);
  Fr<R> ks = sph(kr);
  V3<R> Ehk = ray_to_norm(V3<R>{ks.sp, -ks.cp, R(0)}, f);
  V3<R> Evk = ray_to_norm(V3<R>{ks.ct * ks.cp, ks.ct * ks.sp, -ks.st}, f);
  R cpsi = Eh.x * Ehk.x + Eh.y * Ehk.y;
  R spsi = Eh.x * Evk.x + Eh.y * Evk.y;
  Cx<R> S11, S22;
  amplitude(s, kr.z, S11, S22);
  // Rψ·diag(S11,S22)·Rϕ with Rϕ = [[c, s], [-s, c]], Rψ = [[cψ, sψ], [-sψ, cψ]]
  R c = ks.cp, sn = ks.sp;
  Lmap<R> L;
  L.a = S11 * (cpsi * c) + S22 * (spsi * -sn);
  L.b = S11 * (cpsi * sn) + S22 * (spsi * c);
  L.c = S11 * (-spsi * c) + S22 * (cpsi * -sn);
  L.d = S11 * (-spsi * sn) + S22 * (cpsi * c);
  return L;
}

template <typename R> MCRT_DEV Lmap<R> mul(const Lmap<R>& X, const Lmap<R>& Y) {
  return {X.a * Y.a + X.b * Y.c, X.a * Y.b + X.b * Y.d, X.c * Y.a + X.d * Y.c,
          X.c * Y.b + X.d * Y.d};
}

// kepler_solver (utils.jl:243-261) with an iteration cap.
template <typename R> MCRT_DEV R kepler(R Mv, R e) {
  R MM = Mv - M<R>::floor(Mv / M<R>::two_pi) * M<R>::two_pi;
  R ea = Mv + R(0.85) * e * jl_sign(M<R>::sin(MM));
  R dea = R(1);
  for (int it = 0; it < 64 && M<R>::fabs(dea) > M<R>::kepler_tol; ++it) {
    R f3 = e * M<R>::cos(ea);
    R f2 = e * M<R>::sin(ea);
    R f1 = R(1) - f3;
    R f = ea - f2 - Mv;
    dea = -f / f1;
    dea = -f / (f1 + f2 * dea / R(2));
    dea = -f / (f1 + f2 * dea / R(2) + f3 * (dea * dea) / R(6));
    ea += dea;
  }
  return ea;
}

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon et al., SC'11).  Key = seed; counter = (ray lo,
// ray hi, draw index, instance).
MCRT_DEV void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                            uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
    uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

struct Rng {
  uint64_t seed;
  uint64_t ray;
  uint32_t instance;
  uint32_t draw;
  MCRT_DEV uint64_t bits() {
    uint32_t o[4];
    philox4x32_10((uint32_t)ray, (uint32_t)(ray >> 32), draw++, instance, (uint32_t)seed,
                  (uint32_t)(seed >> 32), o);
    return ((uint64_t)o[0] << 32) | (uint64_t)o[1];
  }
  // (k + 1/2)·2^-52 with the top 52 bits: in (0, 1).
  MCRT_DEV double u64() { return ((double)(bits() >> 12) + 0.5) * 0x1p-52; }
  // (k + 1/2)·2^-23 with the top 23 bits of the same draw: in (0, 1).
  MCRT_DEV float u32() { return ((float)(uint32_t)(bits() >> 41) + 0.5f) * 0x1p-23f; }
  template <typename R> MCRT_DEV R u();
};
template <> MCRT_DEV double Rng::u<double>() { return u64(); }
template <> MCRT_DEV float Rng::u<float>() { return u32(); }

}  // namespace mcrt
