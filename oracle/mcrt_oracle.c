/*
 * mcrt_oracle.c — TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * A plain-C FP64 restatement of the reference's per-ray Monte Carlo
 * polarized RT/CB transport, following the Julia source operation for
 * operation (same expression order, no FMA contraction: build with
 * -ffp-contract=off).  Each function cites the reference lines it restates.
 *
 * Deliberate, documented differences from the reference:
 *  - rand() is replaced by a counter-based Philox4x32-10 stream (Salmon et
 *    al., SC'11): key = seed, counter = (ray lo, ray hi, draw index, instance).
 *    The uniform is (k + 0.5)·2^-52 with k the top 52 bits of two output
 *    words, so it lies in (0,1) and -log(u) is always finite (the reference's
 *    rand() ∈ [0,1) can give -log(0) = Inf, src/cpu.jl:199).
 *  - acos/sqrt arguments that rounding pushes past their domain (|μ| > 1 by an
 *    ulp) are clamped; the reference would throw a DomainError there.
 *  - the Kepler iteration and rejection loops carry very high iteration caps
 *    (never reached in practice); the reference loops without bound.
 *  - libm (glibc) and complex division (Smith's algorithm) differ from Julia
 *    Base at the ulp level.
 */
#include "mcrt_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.141592653589793
#define OR_2PI 6.283185307179586  /* Julia 2π (exact 2·Float64(π)) */
#define OR_4PI 12.566370614359172 /* Julia 4π */
#define OR_EPS 2.220446049250313e-16
#define OR_MAXL 64

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11).                                                 */
/* ------------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline uint64_t or_draw_bits(uint64_t seed, uint32_t instance, uint64_t ray, uint32_t draw) {
  uint32_t ctr[4] = {(uint32_t)ray, (uint32_t)(ray >> 32), draw, instance};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  or_philox4x32_10(ctr, key, o);
  return ((uint64_t)o[0] << 32) | (uint64_t)o[1];
}

double or_uniform(uint64_t seed, uint32_t instance, uint64_t ray, uint32_t draw) {
  uint64_t w = or_draw_bits(seed, instance, ray, draw);
  return ((double)(w >> 12) + 0.5) * 0x1p-52;
}

float or_uniform_f32(uint64_t seed, uint32_t instance, uint64_t ray, uint32_t draw) {
  uint64_t w = or_draw_bits(seed, instance, ray, draw);
  return ((float)(uint32_t)(w >> 41) + 0.5f) * 0x1p-23f;
}

typedef struct {
  uint64_t seed, ray;
  uint32_t instance, draw;
} or_rng;

static inline double rnd(or_rng* r) { return or_uniform(r->seed, r->instance, r->ray, r->draw++); }

/* ------------------------------------------------------------------------ */
/* Small helpers                                                             */
/* ------------------------------------------------------------------------ */
static inline double jl_sign(double x) {
  if (x > 0) return 1.0;
  if (x < 0) return -1.0;
  return x; /* ±0 or NaN */
}
static inline double clamp1(double x) { return x > 1.0 ? 1.0 : (x < -1.0 ? -1.0 : x); }
static inline double sqrt0(double x) { return sqrt(x < 0.0 ? 0.0 : x); }

static inline or_cplx cmk(double re, double im) { or_cplx z = {re, im}; return z; }
static inline or_cplx cadd(or_cplx a, or_cplx b) { return cmk(a.re + b.re, a.im + b.im); }
static inline or_cplx csub(or_cplx a, or_cplx b) { return cmk(a.re - b.re, a.im - b.im); }
/* Julia: Complex(ar*br - ai*bi, ar*bi + ai*br) */
static inline or_cplx cmul(or_cplx a, or_cplx b) {
  return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline or_cplx cscale(or_cplx a, double s) { return cmk(a.re * s, a.im * s); }
static inline or_cplx cneg(or_cplx a) { return cmk(-a.re, -a.im); }
static inline double cabs_(or_cplx a) { return hypot(a.re, a.im); }
static inline double cabs2(or_cplx a) { return a.re * a.re + a.im * a.im; }
/* Smith's algorithm for a/b. */
static inline or_cplx cdiv(or_cplx a, or_cplx b) {
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, den = b.re + b.im * r;
    return cmk((a.re + a.im * r) / den, (a.im - a.re * r) / den);
  } else {
    double r = b.re / b.im, den = b.im + b.re * r;
    return cmk((a.re * r + a.im) / den, (a.im * r - a.re) / den);
  }
}
/* Real / complex as Julia: a * inv(z). */
static inline or_cplx rdivc(double a, or_cplx z) {
  or_cplx inv = cdiv(cmk(1.0, 0.0), z);
  return cmk(a * inv.re, a * inv.im);
}

/* ------------------------------------------------------------------------ */
/* Frames — src/utils.jl:9-84                                                */
/* ------------------------------------------------------------------------ */
/* spherical_from_cartesian, utils.jl:9-23 */
void or_spherical_from_cartesian(const double x[3], double o[4]) {
  double s2 = x[0] * x[0] + x[1] * x[1];
  if (s2 < 1e-24) {
    o[0] = jl_sign(x[2]); o[1] = 0.0; o[2] = 1.0; o[3] = 0.0;
  } else {
    o[0] = x[2];
    o[1] = sqrt0(1.0 - x[2] * x[2]);
    double sinv = 1.0 / sqrt(s2);
    o[2] = x[0] * sinv;
    o[3] = x[1] * sinv;
  }
}

/* ray_to_norm, utils.jl:30-34 */
void or_ray_to_norm(const double x[3], const double f[4], double o[3]) {
  double ct = f[0], st = f[1], cp = f[2], sp = f[3];
  double r0 = x[0] * ct * cp - x[1] * sp + x[2] * st * cp;
  double r1 = x[0] * ct * sp + x[1] * cp + x[2] * st * sp;
  double r2 = -x[0] * st + x[2] * ct;
  o[0] = r0; o[1] = r1; o[2] = r2;
}

/* norm_to_ray, utils.jl:41-45 */
void or_norm_to_ray(const double x[3], const double f[4], double o[3]) {
  double ct = f[0], st = f[1], cp = f[2], sp = f[3];
  double r0 = x[0] * ct * cp + x[1] * ct * sp - x[2] * st;
  double r1 = -x[0] * sp + x[1] * cp;
  double r2 = x[0] * st * cp + x[1] * st * sp + x[2] * ct;
  o[0] = r0; o[1] = r1; o[2] = r2;
}

/* get_basis, utils.jl:52-62.  Ev's second component is -k₃·Eₕ₁ exactly as
 * written in the reference (not the orthogonal θ-hat). */
void or_get_basis(const double k[3], double Eh[3], double Ev[3]) {
  if (fabs(k[2]) > 1.0 - OR_EPS) {
    Eh[0] = 0.0; Eh[1] = -1.0; Eh[2] = 0.0;
    Ev[0] = jl_sign(k[2]); Ev[1] = 0.0; Ev[2] = 0.0;
  } else {
    double Ev3 = -sqrt0(1.0 - k[2] * k[2]);
    double Ev3inv = 1.0 / Ev3;
    double Eh1 = k[1] * -Ev3inv;
    double Eh2 = k[0] * Ev3inv;
    Eh[0] = Eh1; Eh[1] = Eh2; Eh[2] = 0.0;
    Ev[0] = -k[2] * Eh2; Ev[1] = -k[2] * Eh1; Ev[2] = Ev3;
  }
}

/* rotate(I, cos2ϕ, sin2ϕ), utils.jl:82-84 */
void or_rotate(const double I[4], double c2, double s2, double o[4]) {
  double a = c2 * I[1] + s2 * I[2];
  double b = -s2 * I[1] + c2 * I[2];
  o[0] = I[0]; o[1] = a; o[2] = b; o[3] = I[3];
}

/* ------------------------------------------------------------------------ */
/* Scatterers — src/scatterers/rayleigh.jl, mie.jl                           */
/* ------------------------------------------------------------------------ */
/* Inverse-CDF sample of cosθ given the uniform u:
 * Rayleigh rayleigh.jl:9-11 (Cardano), Mie mie.jl:47-51 and Spline
 * spline.jl:27-31 (csrn lerp; length(csrn) - 1 = random_points). */
double or_phase_sample(const mcrt_scatterer* s, double u) {
  if (s->kind == MCRT_SCATTERER_RAYLEIGH) {
    double rn = 2.0 * (1.0 - 2.0 * u);
    double q = sqrt(1.0 + rn * rn);
    return cbrt(q + rn) - cbrt(q - rn);
  } else {
    double rn = u * (double)s->random_points;
    long mrn = (long)trunc(rn);
    if (mrn > s->random_points - 1) mrn = s->random_points - 1;
    return ((double)(mrn + 1) - rn) * s->csrn[mrn] + (rn - (double)mrn) * s->csrn[mrn + 1];
  }
}

/* phase_matrix at a given cosθ: Rayleigh rayleigh.jl:12-17, Mie mie.jl:53-71,
 * Spline spline.jl:26-58 with NSPL read as Nθ - 1 (the name is undefined in
 * the reference) and its weights (msn+1-sn, sn-msn) as written.
 * P[0..5] = (P11, P12, P22, P33, P34, P44); P21 = P12, P43 = -P34; Rayleigh
 * and Mie have P22 = P11, P44 = P33. */
void or_phase_eval(const mcrt_scatterer* s, double mu, double P[6]) {
  if (s->kind == MCRT_SCATTERER_RAYLEIGH) {
    double m2 = mu * mu;
    P[0] = 0.75 * (1.0 + m2);
    P[1] = -0.75 * (1.0 - m2);
    P[3] = 1.5 * mu;
    P[4] = 0.0;
    P[2] = P[0];
    P[5] = P[3];
  } else if (s->kind == MCRT_SCATTERER_MIE) {
    int N = s->n_theta;
    double th = acos(clamp1(mu));
    double sn = th / OR_PI * (double)(N - 1);
    long i0 = (long)trunc(sn);
    if (i0 > N - 1) i0 = N - 1;
    if (i0 < 0) i0 = 0;
    double w0 = (double)(i0 + 1) - sn;
    double F11 = w0 * s->F11[i0], F12 = w0 * s->F12[i0];
    double F33 = w0 * s->F33[i0], F34 = w0 * s->F34[i0];
    if (i0 + 1 <= N - 1) {
      double w1 = sn - (double)(i0 + 1) + 1.0;
      F11 += w1 * s->F11[i0 + 1];
      F12 += w1 * s->F12[i0 + 1];
      F33 += w1 * s->F33[i0 + 1];
      F34 += w1 * s->F34[i0 + 1];
    }
    P[0] = F11; P[1] = F12; P[2] = F11; P[3] = F33; P[4] = F34; P[5] = F33;
  } else {
    /* spline.jl:33-50: sn = θ/π·NSPL, msn = trunc(sn) + 1 (1-based), first
     * term (msn+1-sn)·F[msn], second (sn-msn)·F[msn+1] while msn+1 ≤ length */
    int N = s->n_theta;
    double th = acos(clamp1(mu));
    double sn = th / OR_PI * (double)(N - 1);
    long i0 = (long)trunc(sn);
    if (i0 > N - 1) i0 = N - 1;
    if (i0 < 0) i0 = 0;
    const double* T[6] = {s->F11, s->F12, s->F22, s->F33, s->F34, s->F44};
    double w0 = (double)(i0 + 2) - sn;
    for (int c = 0; c < 6; ++c) P[c] = w0 * T[c][i0];
    if (i0 + 1 <= N - 1) {
      double w1 = sn - (double)(i0 + 1);
      for (int c = 0; c < 6; ++c) P[c] += w1 * T[c][i0 + 1];
    }
  }
}

/* amplitude_matrix: Rayleigh rayleigh.jl:20-23; Mie mie.jl:76-92 and Spline
 * spline.jl:60-76 (the same lookup, NSPL read as Nθ - 1) including its
 * one-step-shifted interpolation weights (msn+1-sn, sn-msn). */
void or_amplitude(const mcrt_scatterer* s, double mu, or_cplx S[2]) {
  if (s->kind == MCRT_SCATTERER_RAYLEIGH) {
    S[0] = cmk(0.0, 1.0);
    S[1] = cmk(mu * 0.0, mu * 1.0);
  } else {
    int N = s->n_theta;
    double th = acos(clamp1(mu));
    double sn = th / OR_PI * (double)(N - 1);
    long i0 = (long)trunc(sn);
    if (i0 > N - 1) i0 = N - 1;
    if (i0 < 0) i0 = 0;
    double w0 = (double)(i0 + 2) - sn;
    or_cplx a = cscale(cmk(s->S11[2 * i0], s->S11[2 * i0 + 1]), w0);
    or_cplx b = cscale(cmk(s->S22[2 * i0], s->S22[2 * i0 + 1]), w0);
    if (i0 + 1 <= N - 1) {
      double w1 = sn - (double)(i0 + 1);
      a = cadd(a, cscale(cmk(s->S11[2 * i0 + 2], s->S11[2 * i0 + 3]), w1));
      b = cadd(b, cscale(cmk(s->S22[2 * i0 + 2], s->S22[2 * i0 + 3]), w1));
    }
    S[0] = a; S[1] = b;
  }
}

/* ------------------------------------------------------------------------ */
/* Stokes and field updates — src/utils.jl:91-261                            */
/* ------------------------------------------------------------------------ */
/* update_I, utils.jl:91-112 */
void or_update_I(const mcrt_scatterer* s, const double I[4], const double k[3],
                 const double f[4], double out[4], double Eh[3], double Ev[3]) {
  or_get_basis(k, Eh, Ev);
  double kr[3], ks[4];
  or_norm_to_ray(k, f, kr);
  or_spherical_from_cartesian(kr, ks);
  double ctk = ks[0], stk = ks[1], cpk = ks[2], spk = ks[3];
  double v1[3] = {spk, -cpk, 0.0};
  double v2[3] = {ctk * cpk, ctk * spk, -stk};
  double Ehk[3], Evk[3];
  or_ray_to_norm(v1, f, Ehk);
  or_ray_to_norm(v2, f, Evk);
  double cpsi = Eh[0] * Ehk[0] + Eh[1] * Ehk[1];
  double spsi = -Eh[0] * Evk[0] - Eh[1] * Evk[1];
  double P[6];
  or_phase_eval(s, kr[2], P);
  double Ir[4];
  or_rotate(I, ks[2] * ks[2] - ks[3] * ks[3], -2.0 * ks[3] * ks[2], Ir);
  double Ip[4];
  Ip[0] = P[0] * Ir[0] + P[1] * Ir[1];
  Ip[1] = P[1] * Ir[0] + P[2] * Ir[1];
  Ip[2] = P[3] * Ir[2] + P[4] * Ir[3];
  Ip[3] = -P[4] * Ir[2] + P[5] * Ir[3];
  or_rotate(Ip, cpsi * cpsi - spsi * spsi, 2.0 * spsi * cpsi, out);
}

/* E_from_I, utils.jl:119-150 */
void or_E_from_I(const double I[4], or_cplx E[2]) {
  const double tol = 1e-14;
  double A1, A2, phi2;
  if (I[0] - fabs(I[1]) > tol * I[0]) {
    A1 = sqrt(0.5 * (I[0] - I[1]));
    A2 = sqrt(0.5 * (I[0] + I[1]));
    double q = sqrt(I[0] * I[0] - I[1] * I[1]);
    if (I[2] >= q) {
      phi2 = 0.0;
    } else if (I[2] <= -q) {
      phi2 = OR_PI;
    } else {
      phi2 = acos(I[2] / q);
      if (I[3] > 0.0) phi2 = OR_2PI - phi2;
    }
  } else if (I[0] + I[1] <= tol * I[0]) {
    A1 = sqrt(I[0]); A2 = 0.0; phi2 = 0.0;
  } else {
    A1 = 0.0; A2 = sqrt(I[0]); phi2 = 0.0;
  }
  E[0] = cmk(A1 * 1.0, A1 * 0.0);
  E[1] = cmk(A2 * cos(phi2), A2 * sin(phi2));
}

/* I_from_E, utils.jl:157-163 */
void or_I_from_E(const or_cplx E[2], double o[4]) {
  double a1 = E[0].re, b1 = E[0].im, a2 = E[1].re, b2 = E[1].im;
  double qh = a1 * a1 + b1 * b1;
  double qv = a2 * a2 + b2 * b2;
  /* E₂·conj(E₁) and conj(E₂)·E₁ with Julia's complex product */
  double r1 = a2 * a1 - b2 * (-b1), i1 = a2 * (-b1) + b2 * a1;
  double r2 = a2 * a1 - (-b2) * b1, i2 = a2 * b1 + (-b2) * a1;
  double dre = r1 - r2, dim = i1 - i2;
  o[0] = qh + qv;
  o[1] = -qh + qv;
  o[2] = r1 + r2;
  o[3] = dre * 0.0 - dim * 1.0; /* real((…)·1.0im) */
}

/* update_E, utils.jl:170-192 */
void or_update_E(const mcrt_scatterer* s, const or_cplx E[2], const double k[3],
                 const double f[4], or_cplx out[2], double Eh[3], double Ev[3]) {
  or_get_basis(k, Eh, Ev);
  double kr[3], ks[4];
  or_norm_to_ray(k, f, kr);
  or_spherical_from_cartesian(kr, ks);
  double ctk = ks[0], stk = ks[1], cpk = ks[2], spk = ks[3];
  double v1[3] = {spk, -cpk, 0.0};
  double v2[3] = {ctk * cpk, ctk * spk, -stk};
  double Ehk[3], Evk[3];
  or_ray_to_norm(v1, f, Ehk);
  or_ray_to_norm(v2, f, Evk);
  double cpsi = Eh[0] * Ehk[0] + Eh[1] * Ehk[1];
  double spsi = Eh[0] * Evk[0] + Eh[1] * Evk[1];
  or_cplx S[2];
  or_amplitude(s, kr[2], S);
  or_cplx E1 = cadd(cscale(E[0], cpk), cscale(E[1], spk));
  or_cplx E2 = cadd(cscale(cneg(E[0]), spk), cscale(E[1], cpk));
  or_cplx F1 = cmul(S[0], E1);
  or_cplx F2 = cmul(S[1], E2);
  out[0] = cadd(cscale(F1, cpsi), cscale(F2, spsi));
  out[1] = cadd(cscale(cneg(F1), spsi), cscale(F2, cpsi));
}

static inline double field_norm(const or_cplx E[2]) {
  double a = cabs_(E[0]), b = cabs_(E[1]);
  return sqrt(a * a + b * b);
}

/* forward_E, utils.jl:199-216.  kpath: n 3-vectors (kpath[0] = k at event 1). */
double or_forward_E(const mcrt_scatterer* s, const or_cplx Ein[2], const double* kpath,
                    int64_t n, const double f0[4], or_cplx out[2]) {
  double enorm = 0.0;
  double norm0 = field_norm(Ein);
  or_cplx E[2] = {Ein[0], Ein[1]};
  double ks[4] = {f0[0], f0[1], f0[2], f0[3]};
  for (int64_t j = 1; j < n; ++j) {
    or_cplx En[2];
    double Eh[3], Ev[3];
    or_update_E(s, E, kpath + 3 * j, ks, En, Eh, Ev);
    double norm2 = field_norm(En);
    double r = norm0 / norm2;
    E[0] = cscale(En[0], r);
    E[1] = cscale(En[1], r);
    enorm += log(norm0 / norm2);
    or_spherical_from_cartesian(kpath + 3 * j, ks);
  }
  out[0] = E[0]; out[1] = E[1];
  return enorm;
}

/* backward_E, utils.jl:223-241 */
double or_backward_E(const mcrt_scatterer* s, const or_cplx Ein[2], const double* kpath,
                     int64_t n, const double f0[4], or_cplx out[2]) {
  double enorm = 0.0;
  double norm0 = field_norm(Ein);
  or_cplx E[2] = {Ein[0], Ein[1]};
  double ks[4] = {f0[0], f0[1], f0[2], f0[3]};
  for (int64_t j = n - 1; j >= 1; --j) {
    double mk[3] = {-kpath[3 * j], -kpath[3 * j + 1], -kpath[3 * j + 2]};
    or_cplx En[2];
    double Eh[3], Ev[3];
    or_update_E(s, E, mk, ks, En, Eh, Ev);
    double norm2 = field_norm(En);
    double r = norm0 / norm2;
    E[0] = cscale(En[0], r);
    E[1] = cscale(En[1], r);
    enorm += log(norm0 / norm2);
    or_spherical_from_cartesian(mk, ks);
  }
  out[0] = E[0]; out[1] = E[1];
  return enorm;
}

/* kepler_solver, utils.jl:243-261 (iteration cap added). */
double or_kepler_solver(double M, double e) {
  const double tol = 1e-12;
  double MM = M - floor(M / OR_2PI) * OR_2PI;
  double ea = M + 0.85 * e * jl_sign(sin(MM));
  double dea = 1.0;
  int it = 0;
  while (fabs(dea) > tol && it < 1000) {
    double f3 = e * cos(ea);
    double f2 = e * sin(ea);
    double f1 = 1.0 - f3;
    double f = ea - f2 - M;
    dea = -f / f1;
    dea = -f / (f1 + f2 * dea / 2.0);
    dea = -f / (f1 + f2 * dea / 2.0 + f3 * (dea * dea) / 6.0);
    ea += dea;
    ++it;
  }
  return ea;
}

/* ------------------------------------------------------------------------ */
/* Table builders                                                            */
/* ------------------------------------------------------------------------ */
/* θ_i of range(0.0, Float64(π), N): the correctly rounded i·π/(N-1). */
static double range_theta(int i, int N) {
  if (N == 1) return 0.0;
  if (i == N - 1) return OR_PI;
  return (double)((long double)OR_PI * (long double)i / (long double)(N - 1));
}

/* bhmie, src/scatterers/mie.jl:129-216 */
int or_bhmie(double x, double m_re, double m_im, int Nt, int nextra, int custom_nstop,
             double q[5], double* S1o, double* S2o, double* F11, double* F12, double* F33,
             double* F34) {
  if (!(x > 0) || Nt < 2) return MCRT_ERR_INVALID;
  or_cplx m = cmk(m_re, m_im);
  or_cplx y = cmk(m.re * x, m.im * x);
  int nstop;
  if (custom_nstop == 0) {
    double a = x + 4.0 * cbrt(x) + 2.0;
    double b = x * hypot(m_re, m_im);
    nstop = (int)floor(a > b ? a : b);
  } else {
    nstop = custom_nstop;
  }
  int nmax = nstop + nextra;
  double* mu = (double*)malloc(sizeof(double) * Nt);
  or_cplx* d = (or_cplx*)calloc((size_t)nmax + 1, sizeof(or_cplx)); /* 1-based */
  double* pi_ = (double*)calloc(Nt, sizeof(double));
  double* tau = (double*)calloc(Nt, sizeof(double));
  double* pi0 = (double*)calloc(Nt, sizeof(double));
  double* pi1 = (double*)malloc(sizeof(double) * Nt);
  or_cplx* S1 = (or_cplx*)calloc(Nt, sizeof(or_cplx));
  or_cplx* S2 = (or_cplx*)calloc(Nt, sizeof(or_cplx));
  for (int i = 0; i < Nt; ++i) { mu[i] = cos(range_theta(i, Nt)); pi1[i] = 1.0; }
  for (int n = nmax - 1; n >= 1; --n) {
    or_cplx t = rdivc((double)(n + 1), y);
    d[n] = csub(t, rdivc(1.0, cadd(d[n + 1], t)));
  }
  double psi0 = cos(x), psi1 = sin(x);
  double chi0 = -sin(x), chi1 = cos(x);
  or_cplx xi1 = cmk(psi1, -chi1);
  or_cplx a = cmk(0, 0), b = cmk(0, 0);
  double qsca = 0.0, asymm = 0.0;
  int half = (Nt + 1) / 2;
  for (int n = 1; n <= nstop; ++n) {
    double fn = (double)(2 * n + 1) / (double)(n * (n + 1));
    double psi = (double)(2 * n - 1) * psi1 / x - psi0;
    double chi = (double)(2 * n - 1) * chi1 / x - chi0;
    or_cplx xi = cmk(psi, -chi);
    or_cplx dm = cdiv(d[n], m);
    or_cplx ta = cadd(dm, cmk((double)n / x, 0.0));
    or_cplx an = csub(cscale(ta, psi), cmk(psi1, 0.0));
    an = cdiv(an, csub(cmul(ta, xi), xi1));
    or_cplx dmm = cmul(d[n], m);
    or_cplx tb = cadd(dmm, cmk((double)n / x, 0.0));
    or_cplx bn = cdiv(csub(cscale(tb, psi), cmk(psi1, 0.0)), csub(cmul(tb, xi), xi1));
    double aa = cabs_(an), bb = cabs_(bn);
    qsca += (aa * aa + bb * bb) * (double)(2 * n + 1);
    asymm += (an.re * bn.re + an.im * bn.im) * fn;
    if (n > 1) {
      asymm += (a.re * an.re + a.im * an.im + b.re * bn.re + b.im * bn.im) * (double)(n - 1) *
               (double)(n + 1) / (double)n;
    }
    for (int i = 0; i < half; ++i) {
      int ii = Nt - 1 - i;
      pi_[i] = pi1[i];
      tau[i] = (double)n * mu[i] * pi_[i] - (double)(n + 1) * pi0[i];
      S1[i] = cadd(S1[i], cscale(cadd(cscale(an, pi_[i]), cscale(bn, tau[i])), fn));
      S2[i] = cadd(S2[i], cscale(cadd(cscale(an, tau[i]), cscale(bn, pi_[i])), fn));
      if (i < ii) {
        double p = (n & 1) == 0 ? -1.0 : 1.0;
        S1[ii] = cadd(S1[ii], cscale(cscale(csub(cscale(an, pi_[i]), cscale(bn, tau[i])), fn), p));
        S2[ii] = cadd(S2[ii],
                      cscale(cscale(cadd(cscale(cneg(an), tau[i]), cscale(bn, pi_[i])), fn), p));
      }
    }
    psi0 = psi1; psi1 = psi;
    chi0 = chi1; chi1 = chi;
    xi1 = cmk(psi1, -chi1);
    a = an; b = bn;
    for (int i = 0; i < half; ++i) {
      pi1[i] = ((double)(2 * n + 1) * mu[i] * pi_[i] - (double)(n + 1) * pi0[i]) / (double)n;
      pi0[i] = pi_[i];
    }
  }
  asymm *= 2.0 / qsca;
  qsca *= 2.0 / (x * x);
  double qext = 4.0 / (x * x) * S1[0].re;
  double qabs = qext - qsca;
  double s1e = cabs_(S1[Nt - 1]);
  double qback = 4.0 / (x * x) * (s1e * s1e);
  double coeff = 4.0 / (qsca * (x * x));
  for (int i = 0; i < Nt; ++i) {
    double a1 = cabs2(S1[i]), a2 = cabs2(S2[i]);
    double r = S2[i].re * S1[i].re - S2[i].im * (-S1[i].im);
    double im = S2[i].re * (-S1[i].im) + S2[i].im * S1[i].re;
    if (F11) F11[i] = (a1 + a2) * 0.5 * coeff;
    if (F12) F12[i] = -(a1 - a2) * 0.5 * coeff;
    if (F33) F33[i] = r * coeff;
    if (F34) F34[i] = im * coeff;
    if (S1o) { S1o[2 * i] = S1[i].re; S1o[2 * i + 1] = S1[i].im; }
    if (S2o) { S2o[2 * i] = S2[i].re; S2o[2 * i + 1] = S2[i].im; }
  }
  if (q) { q[0] = qext; q[1] = qsca; q[2] = qabs; q[3] = qback; q[4] = asymm; }
  free(mu); free(d); free(pi_); free(tau); free(pi0); free(pi1); free(S1); free(S2);
  return MCRT_OK;
}

/* linear_interpolation(range(0, π, N), F) evaluated at x ∈ [0, π]. */
static double itp_range(const double* F, int N, double x) {
  double step = OR_PI / (double)(N - 1);
  double u = x / step;
  if (u < 0) u = 0;
  long i = (long)floor(u);
  if (i > N - 2) i = N - 2;
  double d = u - (double)i;
  return (1.0 - d) * F[i] + d * F[i + 1];
}

/* calculate_csrn, src/utils.jl:263-287 */
int or_calculate_csrn(const double* F11, int N, int n, double* csrn) {
  if (N < 2 || n < 1) return MCRT_ERR_INVALID;
  double* xv = (double*)malloc(sizeof(double) * (n + 1));
  double* yv = (double*)malloc(sizeof(double) * (n + 1));
  double npi = (double)n / OR_PI;
  double acc = 0.0;
  for (int i = 0; i <= n; ++i) {
    xv[i] = (double)i / npi;
    acc += itp_range(F11, N, xv[i]) * sin(xv[i]);
    yv[i] = acc;
  }
  double yend = yv[n];
  for (int i = 0; i <= n; ++i) {
    xv[i] = OR_PI - xv[i];
    yv[i] = 1.0 - yv[i] / yend;
  }
  for (int i = 0; i <= n; ++i) {
    double t = xv[i];
    xv[i] = yv[i];
    yv[i] = -cos(t);
  }
  /* reverse, then drop the first knot: knots K[j] = xv[n-1-j], j = 0..n-1 */
  int nk = n;
  double* K = (double*)malloc(sizeof(double) * nk);
  double* V = (double*)malloc(sizeof(double) * nk);
  for (int j = 0; j < nk; ++j) { K[j] = xv[n - 1 - j]; V[j] = yv[n - 1 - j]; }
  for (int i = 0; i <= n; ++i) {
    double qx = (double)i / (double)n;
    if (nk == 1) { csrn[i] = V[0]; continue; }
    /* searchsortedlast, clamped to an interior interval */
    int lo = 0, hi = nk - 1;
    if (qx <= K[0]) {
      lo = 0;
    } else if (qx >= K[nk - 1]) {
      lo = nk - 2;
    } else {
      while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (K[mid] <= qx) lo = mid; else hi = mid;
      }
    }
    double w = (qx - K[lo]) / (K[lo + 1] - K[lo]);
    if (qx < K[0]) w = 0.0;
    if (qx > K[nk - 1]) w = 1.0;
    csrn[i] = (1.0 - w) * V[lo] + w * V[lo + 1];
  }
  free(xv); free(yv); free(K); free(V);
  return MCRT_OK;
}

/* Gauss–Legendre nodes/weights (FastGaussQuadrature.gausslegendre): Newton
 * iteration on the three-term Legendre recurrence, ascending nodes. */
int or_gausslegendre(int n, double* x, double* w) {
  if (n < 1) return MCRT_ERR_INVALID;
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = cos(OR_PI * ((double)i + 0.75) / ((double)n + 0.5));
    double pp = 0.0, p1 = 0.0;
    for (int it = 0; it < 100; ++it) {
      p1 = 1.0;
      double p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      double z1 = z;
      z = z1 - p1 / pp;
      if (fabs(z - z1) <= 1e-16) break;
    }
    /* final derivative at the converged node */
    p1 = 1.0;
    double p2 = 0.0;
    for (int j = 1; j <= n; ++j) {
      double p3 = p2;
      p2 = p1;
      p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
    }
    pp = n * (z * p1 - p2) / (z * z - 1.0);
    double wi = 2.0 / ((1.0 - z * z) * pp * pp);
    x[i] = -z; x[n - 1 - i] = z;
    w[i] = wi; w[n - 1 - i] = wi;
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
  return MCRT_OK;
}

/* initial_weights: plane plane.jl:131-142, sphere sphere.jl:85-90 */
int or_initial_weights(int kind, int Nt, int Np, double* ct, double* st, double* inorm) {
  if (Nt < 1 || Np < 1) return MCRT_ERR_INVALID;
  if (kind == MCRT_GEOMETRY_PLANE) {
    int h = Nt / 2;
    for (int i = 0; i < Nt; ++i) { ct[i] = 0.0; inorm[i] = 0.0; }
    if (h > 0) {
      double* x = (double*)malloc(sizeof(double) * h);
      double* w = (double*)malloc(sizeof(double) * h);
      or_gausslegendre(h, x, w);
      for (int i = 0; i < h; ++i) { ct[i] = (x[i] - 1.0) / 2.0; inorm[i] = w[i] / 2.0; }
      for (int i = 0; i < h; ++i) { ct[h + i] = -ct[h - 1 - i]; inorm[h + i] = inorm[h - 1 - i]; }
      free(x); free(w);
    }
    double s = 1.0 / (2.0 * Np);
    for (int i = 0; i < Nt; ++i) inorm[i] *= s;
  } else {
    double* w = (double*)malloc(sizeof(double) * Nt);
    or_gausslegendre(Nt, ct, w);
    for (int i = 0; i < Nt; ++i) inorm[i] = w[i] / (2.0 * Np);
    free(w);
  }
  for (int i = 0; i < Nt; ++i) st[i] = sqrt(1.0 - ct[i] * ct[i]);
  return MCRT_OK;
}

/* ------------------------------------------------------------------------ */
/* Geometry — src/geometries/plane.jl, sphere.jl                             */
/* ------------------------------------------------------------------------ */
typedef struct {
  int kind, nl;
  double tau[OR_MAXL], omega[OR_MAXL];
  int isinf[OR_MAXL];
  const mcrt_scatterer* sca[OR_MAXL];
} geo_t;

static int geo_init(const mcrt_geometry* g, geo_t* G) {
  memset(G, 0, sizeof(*G));
  G->kind = g->kind;
  if (g->kind == MCRT_GEOMETRY_PLANE) {
    /* validate, plane.jl:52-62 */
    if (!(OR_PI / 2 < g->theta && g->theta <= OR_PI)) return MCRT_ERR_INVALID;
    if (!(0 <= g->phi && g->phi < OR_2PI)) return MCRT_ERR_INVALID;
    if (g->n_layers < 1 || g->n_layers > OR_MAXL || !g->layers) return MCRT_ERR_INVALID;
    G->nl = g->n_layers;
    for (int i = 0; i < g->n_layers; ++i) {
      G->tau[i] = g->layers[i].tau;
      G->omega[i] = g->layers[i].omega;
      G->isinf[i] = g->layers[i].d < 0;
      G->sca[i] = g->layers[i].scatterer;
      if (!G->sca[i]) return MCRT_ERR_INVALID;
      if (i != g->n_layers - 1 && G->isinf[i]) return MCRT_ERR_INVALID;
    }
  } else if (g->kind == MCRT_GEOMETRY_SPHERE) {
    if (!(g->R > 0)) return MCRT_ERR_INVALID; /* sphere.jl:35-37 */
    if (!g->scatterer) return MCRT_ERR_INVALID;
    G->nl = 1;
    G->tau[0] = g->tau;
    G->omega[0] = g->omega;
    G->sca[0] = g->scatterer;
  } else {
    return MCRT_ERR_INVALID;
  }
  return MCRT_OK;
}

static inline double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* distance_to_boundary: plane plane.jl:97-105, sphere sphere.jl:69-73 */
static double g_distance(const geo_t* G, const double x[3], const double k[3], int idx) {
  if (G->kind == MCRT_GEOMETRY_PLANE) {
    if (k[2] > 0) return -x[2] / k[2];
    if (!G->isinf[idx]) return (x[2] + G->tau[idx]) / -k[2];
    return INFINITY;
  } else {
    double kx = dot3(k, x), xx = dot3(x, x);
    return -kx + sqrt(kx * kx + G->tau[0] * G->tau[0] - xx);
  }
}

double or_distance_to_boundary(const mcrt_geometry* g, const double x[3], const double k[3],
                               int idx) {
  geo_t G;
  if (geo_init(g, &G) != MCRT_OK) return NAN;
  return g_distance(&G, x, k, idx);
}

/* pass_boundary: plane plane.jl:107-115, sphere sphere.jl:75 */
static int g_pass(const geo_t* G, const double x[3], int idx) {
  if (G->kind == MCRT_GEOMETRY_PLANE) {
    if (x[2] > 0) return -1;
    if (!G->isinf[idx] && x[2] < -G->tau[idx]) return 1;
    return 0;
  } else {
    return sqrt(dot3(x, x)) > G->tau[0] ? -2 : 0;
  }
}

/* minimum_distance_to_first_boundary: plane.jl:117, sphere.jl:77 */
static double g_min_first(const geo_t* G, const double x[3], int idx) {
  if (G->kind == MCRT_GEOMETRY_PLANE) return idx == 0 ? fabs(x[2]) : INFINITY;
  return G->tau[0] - sqrt(dot3(x, x));
}

/* minimum_distance_to_last_boundary: plane.jl:121-127, sphere.jl:81 */
static double g_min_last(const geo_t* G, const double x[3], int idx) {
  if (G->kind == MCRT_GEOMETRY_PLANE) {
    if (idx == G->nl - 1 && !G->isinf[idx]) return G->tau[idx] - fabs(x[2]);
    return INFINITY;
  }
  return INFINITY;
}

/* ------------------------------------------------------------------------ */
/* Transport — src/utils.jl:294-323, src/cpu.jl                              */
/* ------------------------------------------------------------------------ */
/* scatter, utils.jl:294-323.  Draws: cosθ sample, then azimuth. */
static void or_scatter(const geo_t* G, const double I[4], const double k[3], int idx,
                       or_rng* rng, double Io[4], double ko[3]) {
  const mcrt_scatterer* s = G->sca[idx];
  double ks[4];
  or_spherical_from_cartesian(k, ks);
  double ct = or_phase_sample(s, rnd(rng));
  double P[6];
  or_phase_eval(s, ct, P);
  double eq = -P[1] * I[1] / (P[0] * I[0]);
  double eu = -P[1] * I[2] / (P[0] * I[0]);
  double ee = sqrt(eq * eq + eu * eu);
  double phi;
  if (ee > 1e-12) {
    double gam = acos(clamp1(eq / ee));
    if (eu < 0) gam = OR_2PI - gam;
    double ma = OR_4PI * rnd(rng) + gam - ee * sin(gam);
    double ea = or_kepler_solver(ma, ee);
    phi = 0.5 * (ea - gam);
  } else {
    phi = OR_2PI * rnd(rng);
  }
  double st = sqrt0(1.0 - ct * ct);
  double v[3] = {st * cos(phi), st * sin(phi), ct};
  or_ray_to_norm(v, ks, ko);
  double Eh[3], Ev[3], Ip[4];
  or_update_I(s, I, ko, ks, Ip, Eh, Ev);
  double r = I[0] / Ip[0];
  for (int c = 0; c < 4; ++c) Io[c] = Ip[c] * r;
}

typedef struct {
  const mcrt_config* cfg;
  const geo_t* G;
  double I0[4], k0[3], f0[4];
  /* scratch path history */
  double* kpath;
  double* xpath;
  int* idxpath;
} tctx_t;

/* collect_rt_cpu!, cpu.jl:6-31.  j1 ∈ [jb, je). */
static double collect_rt(const tctx_t* T, double* Irt, const double I[4], const double k[3],
                         const double x[3], int idx, int jb, int je, uint64_t* nev) {
  const mcrt_config* cfg = T->cfg;
  double ks[4];
  or_spherical_from_cartesian(k, ks);
  double acc = 0.0;
  const mcrt_scatterer* sca = T->G->sca[idx];
  int64_t Nt = cfg->n_theta;
  for (int64_t j2 = 0; j2 < cfg->n_phi; ++j2) {
    for (int j1 = jb; j1 < je; ++j1) {
      double kd[3] = {cfg->sin_theta[j1] * cfg->cos_phi[j2], cfg->sin_theta[j1] * cfg->sin_phi[j2],
                      cfg->cos_theta[j1]};
      double Ip[4], Eh[3], Ev[3];
      or_update_I(sca, I, kd, ks, Ip, Eh, Ev);
      double t = g_distance(T->G, x, kd, idx);
      ++*nev;
      if (t <= cfg->tau0) {
        double nx = exp(-t);
        double* dst = Irt + 4 * (j1 + Nt * j2);
        double w = cfg->inorm[j1];
        double I2[4];
        for (int c = 0; c < 4; ++c) I2[c] = Ip[c] * w * nx;
        for (int c = 0; c < 4; ++c) dst[c] += I2[c];
        acc += I2[0];
      }
    }
  }
  return acc;
}

/* collect_cb_cpu!, cpu.jl:38-160 */
static void collect_cb(const tctx_t* T, double* Icb, const double I[4], const double k[3],
                       const double x[3], int idx, int64_t n, uint64_t* nev) {
  const mcrt_config* cfg = T->cfg;
  const geo_t* G = T->G;
  int64_t Ntb = cfg->n_theta_b, Npb = cfg->n_phi_b;
  double kas[4];
  or_spherical_from_cartesian(k, kas);
  const double* kp = T->kpath;
  double kf0[3] = {-kp[0], -kp[1], -kp[2]};
  double kfs[4];
  or_spherical_from_cartesian(kf0, kfs);
  const mcrt_scatterer* sca = G->sca[0];
  int idxa = idx;
  if (n > 1) {
    const double* ki = kp;
    const double* xb = T->xpath;
    double kb[3] = {-kp[3], -kp[4], -kp[5]};
    int idxb = T->idxpath[0];
    double kbs[4];
    or_spherical_from_cartesian(kb, kbs);
    double ta = g_distance(G, x, kf0, idxa);
    double tb = g_distance(G, xb, kf0, idxb);
    double tab = ta - tb;
    double Iaf[4], Eh[3], Ev[3];
    or_update_I(sca, I, kf0, kas, Iaf, Eh, Ev);
    or_cplx Ei[2], Ea[2], Eb[2];
    or_E_from_I(T->I0, Ei);
    double enorma = or_forward_E(sca, Ei, kp, n, T->f0, Ea);
    double enormb = or_backward_E(sca, Ei, kp, n, T->f0, Eb);
    double sc = exp(enorma - enormb);
    Eb[0] = cscale(Eb[0], sc);
    Eb[1] = cscale(Eb[1], sc);
    or_cplx Eaf[2];
    or_update_E(sca, Ea, kf0, kas, Eaf, Eh, Ev);
    /* raa = sqrt(Iaf[1]/|Eaf|^2) (cpu.jl:68-69).  Where rounding leaves
     * Iaf[1] <= 0 or the field is zero, the reference would raise DomainError
     * or propagate NaN; here the event's CB contribution is zero. */
    double eaf2 = Eaf[0].re * Eaf[0].re + Eaf[0].im * Eaf[0].im +
                  Eaf[1].re * Eaf[1].re + Eaf[1].im * Eaf[1].im;
    double raa = (Iaf[0] > 0 && eaf2 > 0) ? sqrt(Iaf[0] / eaf2) : 0.0;
    Ea[0] = cscale(Ea[0], raa); Ea[1] = cscale(Ea[1], raa);
    Eb[0] = cscale(Eb[0], raa); Eb[1] = cscale(Eb[1], raa);
    double kl = G->sca[0]->kl; /* mean_free_path(g, 1) */
    double dx[3] = {x[0] - xb[0], x[1] - xb[1], x[2] - xb[2]};
    for (int64_t j2 = 0; j2 < Npb; ++j2) {
      for (int64_t j1 = 0; j1 < Ntb; ++j1) {
        double v[3] = {cfg->sin_theta_b[j1] * cfg->cos_phi_b[j2],
                       cfg->sin_theta_b[j1] * cfg->sin_phi_b[j2], cfg->cos_theta_b[j1]};
        double kf[3];
        or_ray_to_norm(v, kfs, kf);
        or_cplx Eaf2[2], Ebf2[2];
        double Eh2[3], Ev2[3];
        or_update_E(sca, Ea, kf, kas, Eaf2, Eh, Ev);
        or_update_E(sca, Eb, kf, kbs, Ebf2, Eh2, Ev2);
        double kik[3] = {ki[0] + kf[0], ki[1] + kf[1], ki[2] + kf[2]};
        double dphi = kl * dot3(kik, dx);
        double ta2 = g_distance(G, x, kf, idxa);
        ++*nev;
        if (ta2 <= cfg->tau0) {
          double nexta = exp(-0.5 * ta2);
          Eaf2[0] = cscale(Eaf2[0], nexta);
          Eaf2[1] = cscale(Eaf2[1], nexta);
          double Iaf2[4];
          or_I_from_E(Eaf2, Iaf2);
          double tb2 = g_distance(G, xb, kf, idxb);
          if (fabs(tb2 + tab) <= cfg->tau0) {
            double nextb = exp(-0.5 * (tb2 + tab));
            Ebf2[0] = cscale(Ebf2[0], nextb);
            Ebf2[1] = cscale(Ebf2[1], nextb);
          } else {
            Ebf2[0] = cmk(0, 0);
            Ebf2[1] = cmk(0, 0);
          }
          double Ibf[4];
          or_I_from_E(Ebf2, Ibf);
          or_cplx ep = cmk(cos(dphi), sin(dphi));
          or_cplx Eab[2] = {cadd(cmul(Eaf2[0], ep), Ebf2[0]), cadd(cmul(Eaf2[1], ep), Ebf2[1])};
          double Iab[4];
          or_I_from_E(Eab, Iab);
          /* Iaf/(Iaf + Ibf) (cpu.jl:114); 0 where both fields vanish (0/0) */
          double den = Iaf2[0] + Ibf[0];
          double rr = den > 0 ? Iaf2[0] / den : 0.0;
          for (int c = 0; c < 4; ++c) Iab[c] = Iab[c] * rr;
          double Ehr[3], Evr[3];
          or_norm_to_ray(Eh2, kfs, Ehr);
          or_norm_to_ray(Ev2, kfs, Evr);
          double cpsi = cfg->sin_phi_b[j2] * Ehr[0] - cfg->cos_phi_b[j2] * Ehr[1];
          double spsi = -cfg->sin_phi_b[j2] * Evr[0] + cfg->cos_phi_b[j2] * Evr[1];
          double o[4];
          or_rotate(Iab, cpsi * cpsi - spsi * spsi, 2.0 * cpsi * spsi, o);
          double* dst = Icb + 4 * (j1 + Ntb * j2);
          for (int c = 0; c < 4; ++c) dst[c] += o[c];
        }
      }
    }
  } else {
    for (int64_t j2 = 0; j2 < Npb; ++j2) {
      for (int64_t j1 = 0; j1 < Ntb; ++j1) {
        double v[3] = {cfg->sin_theta_b[j1] * cfg->cos_phi_b[j2],
                       cfg->sin_theta_b[j1] * cfg->sin_phi_b[j2], cfg->cos_theta_b[j1]};
        double kf[3];
        or_ray_to_norm(v, kfs, kf);
        double Iaf[4], Eh[3], Ev[3];
        or_update_I(sca, I, kf, kas, Iaf, Eh, Ev);
        double ta = g_distance(G, x, kf, idxa);
        ++*nev;
        ++nev[3]; /* nev = &cnt[2], so this is cnt[5]: single-scattering CB evaluations */
        if (ta <= cfg->tau0) {
          double nexta = exp(-ta);
          for (int c = 0; c < 4; ++c) Iaf[c] = Iaf[c] * nexta;
          double Ehr[3], Evr[3];
          or_norm_to_ray(Eh, kfs, Ehr);
          or_norm_to_ray(Ev, kfs, Evr);
          double cpsi = cfg->sin_phi_b[j2] * Ehr[0] - cfg->cos_phi_b[j2] * Ehr[1];
          double spsi = -cfg->sin_phi_b[j2] * Evr[0] + cfg->cos_phi_b[j2] * Evr[1];
          double o[4];
          or_rotate(Iaf, cpsi * cpsi - spsi * spsi, 2.0 * cpsi * spsi, o);
          double* dst = Icb + 4 * (j1 + Ntb * j2);
          for (int c = 0; c < 4; ++c) dst[c] += o[c];
        }
      }
    }
  }
}

/* initial_propagation: plane plane.jl:77-95, sphere sphere.jl:54-67 */
static void initial_propagation(const tctx_t* T, double x[3], const double k[3], or_rng* rng) {
  const geo_t* G = T->G;
  const mcrt_config* cfg = T->cfg;
  double t;
  if (G->kind == MCRT_GEOMETRY_PLANE) {
    if (G->isinf[0]) {
      t = -log(rnd(rng));
    } else {
      int has_dt = (G->nl == 1 && !G->isinf[0]);
      double tp = g_distance(G, x, k, 0);
      if (has_dt && tp <= cfg->tau0) {
        t = -log(1.0 - rnd(rng) * (1.0 - exp(-tp)));
      } else {
        t = tp + 1.0;
        int it = 0;
        while (t > tp && it++ < 100000000) t = -log(rnd(rng));
      }
    }
  } else {
    double tp = 2.0 * x[2];
    if (tp <= cfg->tau0) {
      t = -log(1.0 - rnd(rng) * (1.0 - exp(-tp)));
    } else {
      t = tp + 1.0;
      int it = 0;
      while (t > tp && it++ < 100000000) t = -log(1.0 - rnd(rng));
    }
  }
  x[0] = x[0] + t * k[0];
  x[1] = x[1] + t * k[1];
  x[2] = x[2] + t * k[2];
}

/* cpu_kernel, cpu.jl:162-273 */
static void trace_one(const tctx_t* T, or_rng* rng, double A[5], double* Irt, double* Icb,
                      uint64_t cnt[6]) {
  const mcrt_config* cfg = T->cfg;
  const geo_t* G = T->G;
  int Nt = (int)cfg->n_theta;
  int upper_b, upper_e, lower_b, lower_e;
  if (G->kind == MCRT_GEOMETRY_PLANE) {
    upper_b = Nt / 2; upper_e = Nt; lower_b = 0; lower_e = Nt / 2;
  } else {
    upper_b = 0; upper_e = Nt; lower_b = 0; lower_e = 0;
  }
  double I[4] = {T->I0[0], T->I0[1], T->I0[2], T->I0[3]};
  double k[3] = {T->k0[0], T->k0[1], T->k0[2]};
  double x[3];
  if (G->kind == MCRT_GEOMETRY_PLANE) {
    x[0] = 0.0; x[1] = 0.0; x[2] = 0.0;
  } else { /* initial_x, sphere.jl:43-47 */
    double r = G->tau[0] * sqrt(rnd(rng));
    double ph = OR_2PI * rnd(rng);
    x[0] = r * cos(ph); x[1] = r * sin(ph);
    x[2] = sqrt(G->tau[0] * G->tau[0] - r * r);
  }
  int idx = 0;
  int64_t n = 0;
  int can_dt = (G->kind == MCRT_GEOMETRY_SPHERE) || (G->nl == 1 && !G->isinf[0]);
  if (can_dt) {
    double tp = g_distance(G, x, k, idx);
    double dt = tp <= cfg->tau0 ? exp(-tp) : 0.0;
    A[0] += I[0] * dt;
    for (int c = 0; c < 4; ++c) I[c] = I[c] * (1.0 - dt);
  }
  initial_propagation(T, x, k, rng);
  while (I[0] >= cfg->minimum_intensity && n < cfg->maximum_scattering_times) {
    n += 1;
    memcpy(T->kpath + 3 * (n - 1), k, sizeof(double) * 3);
    memcpy(T->xpath + 3 * (n - 1), x, sizeof(double) * 3);
    T->idxpath[n - 1] = idx;
    double Ip[4], kp[3], xp[3];
    int idxp = idx;
    for (long it = 0;; ++it) {
      or_scatter(G, I, k, idx, rng, Ip, kp);
      double t = -log(rnd(rng));
      xp[0] = x[0] + t * kp[0]; xp[1] = x[1] + t * kp[1]; xp[2] = x[2] + t * kp[2];
      int pb = g_pass(G, xp, idx);
      idxp = idx + pb;
      int ok = 0;
      if (pb == -2) {
        ok = 0;
      } else if (pb == -1) {
        if (idxp < 0) ok = 0;
        else { xp[0] = x[0]; xp[1] = x[1]; xp[2] = -G->tau[idx]; ok = 1; }
      } else if (pb == 1) {
        if (idxp > G->nl - 1) ok = 0;
        else { xp[0] = x[0]; xp[1] = x[1]; xp[2] = 0.0; ok = 1; }
      } else {
        ok = 1;
      }
      if (ok) break;
      cnt[3] += 1;
      if (it > 100000000) { idxp = idx; memcpy(xp, x, sizeof xp); break; }
    }
    /* Absorption */
    double om = G->omega[idx];
    A[1] += (1.0 - om) * I[0];
    for (int c = 0; c < 4; ++c) I[c] = I[c] * om;
    double Iref = 0.0, Itrans = 0.0;
    int near_upper = g_min_first(G, x, idx) <= cfg->tau0;
    int near_lower = g_min_last(G, x, idx) <= cfg->tau0;
    if (cfg->compute_cb && (near_upper || near_lower)) collect_cb(T, Icb, I, k, x, idx, n, &cnt[2]);
    if (near_upper) Iref = collect_rt(T, Irt, I, k, x, idx, upper_b, upper_e, &cnt[1]);
    if (near_lower) Itrans = collect_rt(T, Irt, I, k, x, idx, lower_b, lower_e, &cnt[1]);
    A[2] += Iref;
    A[3] += Itrans;
    double Iesc = Iref + Itrans;
    double s = I[0] - Iesc;
    for (int c = 0; c < 4; ++c) I[c] = Ip[c] * s / Ip[0];
    memcpy(k, kp, sizeof k);
    memcpy(x, xp, sizeof x);
    idx = idxp;
    cnt[0] += 1;
  }
  if (!isfinite(I[0])) cnt[4] += 1;
  A[4] += I[0];
}

static void incident_frame(const mcrt_geometry* g, double k0[3], double f0[4]) {
  double th, ph;
  if (g->kind == MCRT_GEOMETRY_PLANE) { th = g->theta; ph = g->phi; }
  else { th = OR_PI; ph = 0.0; } /* incidence_angle, sphere.jl:41 */
  f0[0] = cos(th); f0[1] = sin(th); f0[2] = cos(ph); f0[3] = sin(ph);
  double z[3] = {0.0, 0.0, 1.0};
  or_ray_to_norm(z, f0, k0); /* simulate, MonteCarloRadiativeTransfer.jl:27-32 */
}

static int check_cfg(const mcrt_config* c) {
  if (!c || c->number_of_rays < 0 || c->maximum_scattering_times < 0) return MCRT_ERR_INVALID;
  if (c->n_theta < 1 || c->n_phi < 1 || !c->cos_theta || !c->sin_theta || !c->inorm ||
      !c->cos_phi || !c->sin_phi)
    return MCRT_ERR_INVALID;
  if (c->compute_cb && (c->n_theta_b < 1 || c->n_phi_b < 1 || !c->cos_theta_b ||
                        !c->sin_theta_b || !c->cos_phi_b || !c->sin_phi_b))
    return MCRT_ERR_INVALID;
  return MCRT_OK;
}

int or_trace_ray(const mcrt_config* cfg, const mcrt_geometry* g, const double I0[4],
                 uint64_t seed, uint32_t instance, uint64_t ray, double A[5], double* Irt,
                 double* Icb, uint64_t counters[6]) {
  geo_t G;
  int rc = geo_init(g, &G);
  if (rc) return rc;
  if ((rc = check_cfg(cfg))) return rc;
  tctx_t T;
  T.cfg = cfg; T.G = &G;
  memcpy(T.I0, I0, sizeof T.I0);
  incident_frame(g, T.k0, T.f0);
  int64_t ms = cfg->maximum_scattering_times > 0 ? cfg->maximum_scattering_times : 1;
  T.kpath = (double*)malloc(sizeof(double) * 3 * ms);
  T.xpath = (double*)malloc(sizeof(double) * 3 * ms);
  T.idxpath = (int*)malloc(sizeof(int) * ms);
  or_rng rng = {seed, ray, instance, 0};
  trace_one(&T, &rng, A, Irt, Icb, counters);
  free(T.kpath); free(T.xpath); free(T.idxpath);
  return MCRT_OK;
}

int or_collect_cb_event(const mcrt_config* cfg, const mcrt_geometry* g, const double I0[4],
                        const double I[4], const double k[3], const double x[3], int idx,
                        int64_t n, const double* kpath, const double* xpath, const int* idxpath,
                        double* Icb) {
  geo_t G;
  int rc = geo_init(g, &G);
  if (rc) return rc;
  if ((rc = check_cfg(cfg))) return rc;
  if (!cfg->compute_cb || n < 1 || !kpath || !xpath || !idxpath || !Icb) return MCRT_ERR_INVALID;
  tctx_t T;
  T.cfg = cfg; T.G = &G;
  memcpy(T.I0, I0, sizeof T.I0);
  incident_frame(g, T.k0, T.f0);
  T.kpath = (double*)kpath;
  T.xpath = (double*)xpath;
  T.idxpath = (int*)idxpath;
  uint64_t cnt[6] = {0, 0, 0, 0, 0, 0};
  collect_cb(&T, Icb, I, k, x, idx, n, &cnt[2]);
  return MCRT_OK;
}

typedef struct {
  const mcrt_config* cfg;
  const geo_t* G;
  const mcrt_geometry* g;
  const double* I0;
  uint64_t seed;
  uint32_t instance;
  int64_t rb, re;
  int nbatch;
  double* batches;
  int64_t stride;
  int next; /* shared batch counter */
  uint64_t cnt[6];
  pthread_mutex_t mu;
} sim_t;

static void* sim_worker(void* arg) {
  sim_t* S = (sim_t*)arg;
  tctx_t T;
  T.cfg = S->cfg; T.G = S->G;
  memcpy(T.I0, S->I0, sizeof T.I0);
  incident_frame(S->g, T.k0, T.f0);
  int64_t ms = S->cfg->maximum_scattering_times > 0 ? S->cfg->maximum_scattering_times : 1;
  T.kpath = (double*)malloc(sizeof(double) * 3 * ms);
  T.xpath = (double*)malloc(sizeof(double) * 3 * ms);
  T.idxpath = (int*)malloc(sizeof(int) * ms);
  uint64_t cnt[6] = {0, 0, 0, 0, 0, 0};
  int64_t nrt = 4 * S->cfg->n_theta * S->cfg->n_phi;
  for (;;) {
    int b = __atomic_fetch_add(&S->next, 1, __ATOMIC_RELAXED);
    if (b >= S->nbatch) break;
    int64_t tot = S->re - S->rb;
    int64_t r0 = S->rb + (tot * b) / S->nbatch;
    int64_t r1 = S->rb + (tot * (b + 1)) / S->nbatch;
    double* out = S->batches + S->stride * b;
    memset(out, 0, sizeof(double) * S->stride);
    for (int64_t r = r0; r < r1; ++r) {
      or_rng rng = {S->seed, (uint64_t)r, S->instance, 0};
      trace_one(&T, &rng, out, out + 5, out + 5 + nrt, cnt);
    }
  }
  pthread_mutex_lock(&S->mu);
  for (int c = 0; c < 6; ++c) S->cnt[c] += cnt[c];
  pthread_mutex_unlock(&S->mu);
  free(T.kpath); free(T.xpath); free(T.idxpath);
  return NULL;
}

int or_simulate(const mcrt_config* cfg, const mcrt_geometry* g, const double I0[4],
                uint64_t seed, uint32_t instance, int64_t ray_begin, int64_t ray_end,
                int nthreads, int nbatch, double* batches, uint64_t counters[6]) {
  geo_t G;
  int rc = geo_init(g, &G);
  if (rc) return rc;
  if ((rc = check_cfg(cfg))) return rc;
  if (nbatch < 1 || ray_end < ray_begin || !batches) return MCRT_ERR_INVALID;
  if (nthreads < 1) nthreads = 1;
  sim_t S;
  memset(&S, 0, sizeof S);
  S.cfg = cfg; S.G = &G; S.g = g; S.I0 = I0; S.seed = seed; S.instance = instance;
  S.rb = ray_begin; S.re = ray_end; S.nbatch = nbatch; S.batches = batches;
  int64_t ncb = cfg->compute_cb || cfg->n_theta_b > 0 ? 4 * cfg->n_theta_b * cfg->n_phi_b : 0;
  S.stride = 5 + 4 * cfg->n_theta * cfg->n_phi + ncb;
  pthread_mutex_init(&S.mu, NULL);
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
  for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, sim_worker, &S);
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  free(th);
  pthread_mutex_destroy(&S.mu);
  if (counters) for (int c = 0; c < 6; ++c) counters[c] = S.cnt[c];
  return MCRT_OK;
}
