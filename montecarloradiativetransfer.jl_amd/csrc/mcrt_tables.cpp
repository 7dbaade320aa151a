// mcrt_tables.cpp — host FP64 scatterer-table and quadrature builders of
// libmcrt_hip.so (no device use).
//
//   mcrt_bhmie            src/scatterers/mie.jl:129-216  (Bohren–Huffman series)
//   mcrt_calculate_csrn   src/utils.jl:263-287 with src/scatterers/mie.jl:41-42
//   mcrt_gauss_legendre   FastGaussQuadrature.gausslegendre (plane.jl:132, sphere.jl:86)
//   mcrt_initial_weights  src/geometries/plane.jl:131-142, sphere.jl:85-90
//
// The Mie recurrences stay in FP64 in every precision form of the transport.
// Expression order follows the Julia source, so results agree with it to a
// few ulps (complex division and libm differ from Julia Base at ulp level).
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/mcrt.h"
#include "mcrt_internal.h"

namespace {

constexpr double kPi = 3.141592653589793;

struct C {
  double re, im;
};
inline C operator+(C a, C b) { return {a.re + b.re, a.im + b.im}; }
inline C operator-(C a, C b) { return {a.re - b.re, a.im - b.im}; }
inline C operator-(C a) { return {-a.re, -a.im}; }
// Julia's complex product: (ar*br - ai*bi, ar*bi + ai*br)
inline C operator*(C a, C b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline C operator*(C a, double s) { return {a.re * s, a.im * s}; }
inline C operator*(double s, C a) { return {s * a.re, s * a.im}; }
inline C operator-(C a, double s) { return {a.re - s, a.im}; }
inline C operator+(C a, double s) { return {a.re + s, a.im}; }
// Smith's scaled division.
inline C operator/(C a, C b) {
  if (std::fabs(b.re) >= std::fabs(b.im)) {
    double r = b.im / b.re, den = b.re + b.im * r;
    return {(a.re + a.im * r) / den, (a.im - a.re * r) / den};
  }
  double r = b.re / b.im, den = b.im + b.re * r;
  return {(a.re * r + a.im) / den, (a.im * r - a.re) / den};
}
// real / complex evaluated as a·inv(z), as Julia does.
inline C rdiv(double a, C z) { return a * (C{1.0, 0.0} / z); }
inline double cabs(C a) { return std::hypot(a.re, a.im); }
inline double cabs2(C a) { return a.re * a.re + a.im * a.im; }

// range(0.0, Float64(π), N)[i] (0-based), correctly rounded like Julia's
// TwicePrecision ranges.
double range_theta(int i, int N) {
  if (N == 1) return 0.0;
  if (i == N - 1) return kPi;
  return static_cast<double>(static_cast<long double>(kPi) * i / (N - 1));
}

}  // namespace

extern "C" int mcrt_bhmie(double x, double m_re, double m_im, int32_t n_theta, int32_t nextra,
                          int32_t custom_nstop, double* q, double* S1o, double* S2o,
                          double* F11, double* F12, double* F33, double* F34) {
  if (!(x > 0.0) || !std::isfinite(x)) return mcrt::fail(nullptr, MCRT_ERR_INVALID, "bhmie: x must be positive and finite");
  if (n_theta < 2) return mcrt::fail(nullptr, MCRT_ERR_INVALID, "bhmie: Nθ must be ≥ 2");
  if (nextra < 1 || custom_nstop < 0) return mcrt::fail(nullptr, MCRT_ERR_INVALID, "bhmie: bad nextra/custom_nstop");
  const int N = n_theta;
  const C m{m_re, m_im};
  const C y = m * x;
  int nstop = custom_nstop;
  if (nstop == 0) {
    double a = x + 4.0 * std::cbrt(x) + 2.0;
    double b = x * std::hypot(m_re, m_im);
    nstop = static_cast<int>(std::floor(a > b ? a : b));
  }
  const int nmax = nstop + nextra;
  std::vector<double> mu(N), pic(N, 0.0), tau(N, 0.0), pi0(N, 0.0), pi1(N, 1.0);
  std::vector<C> d(nmax + 1, C{0.0, 0.0}), S1(N, C{0, 0}), S2(N, C{0, 0});
  for (int i = 0; i < N; ++i) mu[i] = std::cos(range_theta(i, N));
  // logarithmic derivative D_n by downward recurrence from D_nmax = 0
  for (int n = nmax - 1; n >= 1; --n) {
    C t = rdiv(static_cast<double>(n + 1), y);
    d[n] = t - rdiv(1.0, d[n + 1] + t);
  }
  double psi0 = std::cos(x), psi1 = std::sin(x);
  double chi0 = -std::sin(x), chi1 = std::cos(x);
  C xi1{psi1, -chi1};
  C ap{0, 0}, bp{0, 0};
  double qsca = 0.0, asymm = 0.0;
  const int half = (N + 1) / 2;
  for (int n = 1; n <= nstop; ++n) {
    const double fn = static_cast<double>(2 * n + 1) / static_cast<double>(n * (n + 1));
    const double psi = static_cast<double>(2 * n - 1) * psi1 / x - psi0;
    const double chi = static_cast<double>(2 * n - 1) * chi1 / x - chi0;
    const C xi{psi, -chi};
    const C ta = d[n] / m + static_cast<double>(n) / x;
    C an = ta * psi - psi1;
    an = an / (ta * xi - xi1);
    const C tb = d[n] * m + static_cast<double>(n) / x;
    const C bn = (tb * psi - psi1) / (tb * xi - xi1);
    const double aa = cabs(an), bb = cabs(bn);
    qsca += (aa * aa + bb * bb) * static_cast<double>(2 * n + 1);
    asymm += (an.re * bn.re + an.im * bn.im) * fn;
    if (n > 1)
      asymm += (ap.re * an.re + ap.im * an.im + bp.re * bn.re + bp.im * bn.im) *
               static_cast<double>(n - 1) * static_cast<double>(n + 1) / static_cast<double>(n);
    const double par = (n & 1) == 0 ? -1.0 : 1.0;
    for (int i = 0; i < half; ++i) {
      const int ii = N - 1 - i;
      pic[i] = pi1[i];
      tau[i] = static_cast<double>(n) * mu[i] * pic[i] - static_cast<double>(n + 1) * pi0[i];
      S1[i] = S1[i] + fn * (an * pic[i] + bn * tau[i]);
      S2[i] = S2[i] + fn * (an * tau[i] + bn * pic[i]);
      if (i < ii) {
        S1[ii] = S1[ii] + fn * (an * pic[i] - bn * tau[i]) * par;
        S2[ii] = S2[ii] + fn * (-an * tau[i] + bn * pic[i]) * par;
      }
    }
    psi0 = psi1; psi1 = psi;
    chi0 = chi1; chi1 = chi;
    xi1 = C{psi1, -chi1};
    ap = an; bp = bn;
    for (int i = 0; i < half; ++i) {
      pi1[i] = (static_cast<double>(2 * n + 1) * mu[i] * pic[i] -
                static_cast<double>(n + 1) * pi0[i]) / static_cast<double>(n);
      pi0[i] = pic[i];
    }
  }
  asymm *= 2.0 / qsca;
  qsca *= 2.0 / (x * x);
  const double qext = 4.0 / (x * x) * S1[0].re;
  const double sb = cabs(S1[N - 1]);
  const double qback = 4.0 / (x * x) * (sb * sb);
  const double coeff = 4.0 / (qsca * (x * x));
  for (int i = 0; i < N; ++i) {
    const double a1 = cabs2(S1[i]), a2 = cabs2(S2[i]);
    const C s21 = S2[i] * C{S1[i].re, -S1[i].im};
    if (F11) F11[i] = (a1 + a2) * 0.5 * coeff;
    if (F12) F12[i] = -(a1 - a2) * 0.5 * coeff;
    if (F33) F33[i] = s21.re * coeff;
    if (F34) F34[i] = s21.im * coeff;
    if (S1o) { S1o[2 * i] = S1[i].re; S1o[2 * i + 1] = S1[i].im; }
    if (S2o) { S2o[2 * i] = S2[i].re; S2o[2 * i + 1] = S2[i].im; }
  }
  if (q) { q[0] = qext; q[1] = qsca; q[2] = qext - qsca; q[3] = qback; q[4] = asymm; }
  return MCRT_OK;
}

extern "C" int mcrt_calculate_csrn(const double* F11, int32_t N, int32_t n, double* csrn) {
  if (!F11 || !csrn || N < 2 || n < 1)
    return mcrt::fail(nullptr, MCRT_ERR_INVALID, "calculate_csrn: bad arguments");
  // itp11 = linear_interpolation(range(0, π, N), F11)
  const double step = kPi / static_cast<double>(N - 1);
  auto itp11 = [&](double t) {
    double u = t / step;
    if (u < 0.0) u = 0.0;
    long i = static_cast<long>(std::floor(u));
    if (i > N - 2) i = N - 2;
    double w = u - static_cast<double>(i);
    return (1.0 - w) * F11[i] + w * F11[i + 1];
  };
  std::vector<double> xv(n + 1), yv(n + 1);
  const double npi = static_cast<double>(n) / kPi;
  double run = 0.0;
  for (int i = 0; i <= n; ++i) {  // rectangle-rule CDF of F11(θ)·sinθ
    xv[i] = static_cast<double>(i) / npi;
    run += itp11(xv[i]) * std::sin(xv[i]);
    yv[i] = run;
  }
  const double total = yv[n];
  // swap axes to (1 - CDF, cos θ) after θ → π - θ, reverse, drop the first knot
  std::vector<double> K(n), V(n);
  for (int j = 0; j < n; ++j) {
    const int i = n - 1 - j;
    K[j] = 1.0 - yv[i] / total;
    V[j] = -std::cos(kPi - xv[i]);
  }
  for (int i = 0; i <= n; ++i) {
    const double u = static_cast<double>(i) / static_cast<double>(n);
    if (n == 1) { csrn[i] = V[0]; continue; }
    int lo;
    if (u <= K[0]) {
      lo = 0;
    } else if (u >= K[n - 1]) {
      lo = n - 2;
    } else {
      int a = 0, b = n - 1;
      while (b - a > 1) {
        const int c = (a + b) / 2;
        if (K[c] <= u) a = c; else b = c;
      }
      lo = a;
    }
    double w = (u - K[lo]) / (K[lo + 1] - K[lo]);
    if (u < K[0]) w = 0.0;
    if (u > K[n - 1]) w = 1.0;
    csrn[i] = (1.0 - w) * V[lo] + w * V[lo + 1];
  }
  return MCRT_OK;
}

extern "C" int mcrt_gauss_legendre(int32_t n, double* x, double* w) {
  if (n < 1 || !x || !w) return mcrt::fail(nullptr, MCRT_ERR_INVALID, "gauss_legendre: n must be ≥ 1");
  // Newton's method on P_n from the three-term recurrence, started at the
  // Tricomi-type guess cos(π(i + 3/4)/(n + 1/2)).
  auto legendre = [n](double z, double& pn, double& dpn) {
    double p0 = 1.0, pm = 0.0;
    for (int j = 1; j <= n; ++j) {
      const double pmm = pm;
      pm = p0;
      p0 = ((2.0 * j - 1.0) * z * pm - (j - 1.0) * pmm) / j;
    }
    pn = p0;
    dpn = n * (z * p0 - pm) / (z * z - 1.0);
  };
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
    double pn, dpn;
    for (int it = 0; it < 100; ++it) {
      legendre(z, pn, dpn);
      const double dz = pn / dpn;
      z -= dz;
      if (std::fabs(dz) <= 1e-16) break;
    }
    legendre(z, pn, dpn);
    const double wi = 2.0 / ((1.0 - z * z) * dpn * dpn);
    x[i] = -z; x[n - 1 - i] = z;
    w[i] = wi; w[n - 1 - i] = wi;
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
  return MCRT_OK;
}

extern "C" int mcrt_initial_weights(int32_t kind, int32_t Nt, int32_t Np, double* ct, double* st,
                                    double* inorm) {
  if (Nt < 1 || Np < 1 || !ct || !st || !inorm)
    return mcrt::fail(nullptr, MCRT_ERR_INVALID, "initial_weights: bad arguments");
  if (kind == MCRT_GEOMETRY_PLANE) {
    const int h = Nt / 2;
    for (int i = 0; i < Nt; ++i) { ct[i] = 0.0; inorm[i] = 0.0; }
    if (h > 0) {
      std::vector<double> gx(h), gw(h);
      mcrt_gauss_legendre(h, gx.data(), gw.data());
      for (int i = 0; i < h; ++i) { ct[i] = (gx[i] - 1.0) / 2.0; inorm[i] = gw[i] / 2.0; }
      for (int i = 0; i < h; ++i) { ct[h + i] = -ct[h - 1 - i]; inorm[h + i] = inorm[h - 1 - i]; }
    }
    const double s = 1.0 / (2.0 * Np);
    for (int i = 0; i < Nt; ++i) inorm[i] *= s;
  } else if (kind == MCRT_GEOMETRY_SPHERE) {
    std::vector<double> gw(Nt);
    mcrt_gauss_legendre(Nt, ct, gw.data());
    for (int i = 0; i < Nt; ++i) inorm[i] = gw[i] / (2.0 * Np);
  } else {
    return mcrt::fail(nullptr, MCRT_ERR_INVALID, "initial_weights: unknown geometry kind");
  }
  for (int i = 0; i < Nt; ++i) st[i] = std::sqrt(1.0 - ct[i] * ct[i]);
  return MCRT_OK;
}
