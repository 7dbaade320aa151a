// mcrt_plain.hip — the plain FP64 drop-in transport kernel.
//
// Same structure as the reference GPU path (src/gpu.jl:177-296): one ray per
// thread in a grid-stride loop, FP64 throughout, per-thread path history
// (kpath/xpath/idxpath) in HBM, the O(n) forward/backward field re-walk for
// CB, and one FP64 global atomic per Stokes component per bin.  It is the
// baseline the optimised forms are measured against and the GPU counterpart of
// the oracle on the same Philox stream.  Floating-point contraction is off so
// that the arithmetic matches the reference's unfused Julia evaluation.
#include "mcrt_kernels.h"

#pragma clang fp contract(off)

namespace mcrt {
namespace {

using R = double;

struct Ctx {
  const PlainParams* p;
  double* kp;   // this thread's kpath column
  double* xp;
  int8_t* ip;
};

MCRT_DEV V3<R> ld3(const double* a, int64_t s) { return {a[3 * s], a[3 * s + 1], a[3 * s + 2]}; }
MCRT_DEV void st3(double* a, int64_t s, V3<R> v) { a[3 * s] = v.x; a[3 * s + 1] = v.y; a[3 * s + 2] = v.z; }

// distance_to_boundary (plane.jl:97-105, sphere.jl:69-73)
MCRT_DEV R distance(const GeoDev& g, V3<R> x, V3<R> k, int idx) {
  if (g.kind == 0) {
    if (k.z > 0) return -x.z / k.z;
    if (!g.isinf[idx]) return (x.z + g.tau[idx]) / -k.z;
    return __builtin_inf();
  }
  R kx = dot(k, x), xx = dot(x, x);
  return -kx + sqrt(kx * kx + g.tau[0] * g.tau[0] - xx);
}

MCRT_DEV void aadd(double* p, double v) { unsafeAtomicAdd(p, v); }

// collect_rt_gpu! (gpu.jl:6-34)
MCRT_DEV R collect_rt(const Ctx& c, const R I[4], V3<R> k, V3<R> x, int idx, int jb, int je,
                      unsigned long long& nev) {
  const PlainParams& p = *c.p;
  const BinsDev& b = p.bins;
  Fr<R> ks = sph(k);
  R acc = 0.0;
  const ScaView<R>& sca = p.sca[p.geo.sca[idx]];
  for (int j2 = 0; j2 < b.Np; ++j2) {
    for (int j1 = jb; j1 < je; ++j1) {
      V3<R> kd = {b.st[j1] * b.cp[j2], b.st[j1] * b.sp[j2], b.ct[j1]};
      R Ip[4];
      update_I(sca, I, kd, ks, Ip);
      R t = distance(p.geo, x, kd, idx);
      ++nev;
      if (t <= p.tau0) {
        R nx = exp(-t);
        R w = b.inorm[j1];
        double* dst = p.Irt + 4 * (j1 + (int64_t)b.Nt * j2);
        R v0 = Ip[0] * w * nx, v1 = Ip[1] * w * nx, v2 = Ip[2] * w * nx, v3 = Ip[3] * w * nx;
        aadd(dst + 0, v0); aadd(dst + 1, v1); aadd(dst + 2, v2); aadd(dst + 3, v3);
        acc += v0;
      }
    }
  }
  return acc;
}

MCRT_DEV R field_norm(Cx<R> a, Cx<R> b) {
  R x = hypot(a.re, a.im), y = hypot(b.re, b.im);
  return sqrt(x * x + y * y);
}

// collect_cb_gpu! (gpu.jl:41-175), with forward_E/backward_E (utils.jl:199-241)
// re-walking this thread's kpath column.
MCRT_DEV void collect_cb(const Ctx& c, const R I[4], V3<R> k, V3<R> x, int idx, int64_t n,
                         unsigned long long& nev, unsigned long long& nev1) {
  const PlainParams& p = *c.p;
  const BinsDev& b = p.bins;
  const GeoDev& g = p.geo;
  Fr<R> kas = sph(k);
  V3<R> k1 = ld3(c.kp, 0);
  V3<R> kf0 = -k1;
  Fr<R> kfs = sph(kf0);
  const ScaView<R>& sca = p.sca[g.sca[0]];
  Fr<R> f0 = {p.f0[0], p.f0[1], p.f0[2], p.f0[3]};
  if (n > 1) {
    V3<R> ki = k1;
    V3<R> xb = ld3(c.xp, 0);
    V3<R> kb = -ld3(c.kp, 1);
    int idxb = c.ip[0];
    Fr<R> kbs = sph(kb);
    R ta = distance(g, x, kf0, idx);
    R tb = distance(g, xb, kf0, idxb);
    R tab = ta - tb;
    R Iaf[4];
    update_I(sca, I, kf0, kas, Iaf);
    Cx<R> Ei1, Ei2;
    E_from_I(p.I0, Ei1, Ei2);
    // forward_E
    R norm0 = field_norm(Ei1, Ei2);
    Cx<R> Ea1 = Ei1, Ea2 = Ei2;
    R enorma = 0.0;
    Fr<R> ksf = f0;
    for (int64_t j = 1; j < n; ++j) {
      V3<R> kj = ld3(c.kp, j);
      update_E(sca, Ea1, Ea2, kj, ksf);
      R norm2 = field_norm(Ea1, Ea2);
      R r = norm0 / norm2;
      Ea1 = Ea1 * r; Ea2 = Ea2 * r;
      enorma += log(norm0 / norm2);
      ksf = sph(kj);
    }
    // backward_E
    Cx<R> Eb1 = Ei1, Eb2 = Ei2;
    R enormb = 0.0;
    Fr<R> ksb = f0;
    for (int64_t j = n - 1; j >= 1; --j) {
      V3<R> mk = -ld3(c.kp, j);
      update_E(sca, Eb1, Eb2, mk, ksb);
      R norm2 = field_norm(Eb1, Eb2);
      R r = norm0 / norm2;
      Eb1 = Eb1 * r; Eb2 = Eb2 * r;
      enormb += log(norm0 / norm2);
      ksb = sph(mk);
    }
    R sc = exp(enorma - enormb);
    Eb1 = Eb1 * sc; Eb2 = Eb2 * sc;
    Cx<R> Eaf1 = Ea1, Eaf2 = Ea2;
    update_E(sca, Eaf1, Eaf2, kf0, kas);
    // raa (cpu.jl:68-69); zero where rounding leaves Iaf₁ ≤ 0 or the field is
    // zero (the reference would raise DomainError or propagate NaN)
    const R eaf2 = Eaf1.re * Eaf1.re + Eaf1.im * Eaf1.im + Eaf2.re * Eaf2.re + Eaf2.im * Eaf2.im;
    R raa = (Iaf[0] > 0 && eaf2 > 0) ? sqrt(Iaf[0] / eaf2) : R(0);
    Ea1 = Ea1 * raa; Ea2 = Ea2 * raa;
    Eb1 = Eb1 * raa; Eb2 = Eb2 * raa;
    V3<R> dx = x - xb;
    for (int j2 = 0; j2 < b.Npb; ++j2) {
      for (int j1 = 0; j1 < b.Ntb; ++j1) {
        V3<R> v = {b.stb[j1] * b.cpb[j2], b.stb[j1] * b.spb[j2], b.ctb[j1]};
        V3<R> kf = ray_to_norm(v, kfs);
        Cx<R> a1 = Ea1, a2 = Ea2, b1 = Eb1, b2 = Eb2;
        V3<R> Eh, Ev;
        update_E(sca, a1, a2, kf, kas);
        update_E(sca, b1, b2, kf, kbs, &Eh, &Ev);
        R dphi = g.kl1 * dot(ki + kf, dx);
        R ta2 = distance(g, x, kf, idx);
        ++nev;
        if (ta2 <= p.tau0) {
          R na = exp(-0.5 * ta2);
          a1 = a1 * na; a2 = a2 * na;
          R Ia[4];
          I_from_E(a1, a2, Ia);
          R tb2 = distance(g, xb, kf, idxb);
          if (fabs(tb2 + tab) <= p.tau0) {
            R nb = exp(-0.5 * (tb2 + tab));
            b1 = b1 * nb; b2 = b2 * nb;
          } else {
            b1 = {0.0, 0.0}; b2 = {0.0, 0.0};
          }
          R Ib[4];
          I_from_E(b1, b2, Ib);
          R sp_, cp_;
          sincos(dphi, &sp_, &cp_);
          Cx<R> ep = {cp_, sp_};
          Cx<R> e1 = a1 * ep + b1, e2 = a2 * ep + b2;
          R Iab[4];
          I_from_E(e1, e2, Iab);
          const R den = Ia[0] + Ib[0];
          R rr = den > 0 ? Ia[0] / den : R(0);  // 0 where both fields vanish
          for (int q = 0; q < 4; ++q) Iab[q] = Iab[q] * rr;
          V3<R> Ehr = norm_to_ray(Eh, kfs), Evr = norm_to_ray(Ev, kfs);
          R cpsi = b.spb[j2] * Ehr.x - b.cpb[j2] * Ehr.y;
          R spsi = -b.spb[j2] * Evr.x + b.cpb[j2] * Evr.y;
          rotate(Iab, cpsi * cpsi - spsi * spsi, 2.0 * cpsi * spsi);
          double* dst = p.Icb + 4 * (j1 + (int64_t)b.Ntb * j2);
          aadd(dst + 0, Iab[0]); aadd(dst + 1, Iab[1]); aadd(dst + 2, Iab[2]); aadd(dst + 3, Iab[3]);
        }
      }
    }
  } else {
    for (int j2 = 0; j2 < b.Npb; ++j2) {
      for (int j1 = 0; j1 < b.Ntb; ++j1) {
        V3<R> v = {b.stb[j1] * b.cpb[j2], b.stb[j1] * b.spb[j2], b.ctb[j1]};
        V3<R> kf = ray_to_norm(v, kfs);
        R Ia[4];
        V3<R> Eh, Ev;
        update_I(sca, I, kf, kas, Ia, &Eh, &Ev);
        R ta = distance(g, x, kf, idx);
        ++nev;
        ++nev1;
        if (ta <= p.tau0) {
          R na = exp(-ta);
          for (int q = 0; q < 4; ++q) Ia[q] = Ia[q] * na;
          V3<R> Ehr = norm_to_ray(Eh, kfs), Evr = norm_to_ray(Ev, kfs);
          R cpsi = b.spb[j2] * Ehr.x - b.cpb[j2] * Ehr.y;
          R spsi = -b.spb[j2] * Evr.x + b.cpb[j2] * Evr.y;
          rotate(Ia, cpsi * cpsi - spsi * spsi, 2.0 * cpsi * spsi);
          double* dst = p.Icb + 4 * (j1 + (int64_t)b.Ntb * j2);
          aadd(dst + 0, Ia[0]); aadd(dst + 1, Ia[1]); aadd(dst + 2, Ia[2]); aadd(dst + 3, Ia[3]);
        }
      }
    }
  }
}

// scatter (utils.jl:294-323)
MCRT_DEV void scatter(const ScaView<R>& s, const R I[4], V3<R> k, Rng& rng, R Io[4], V3<R>& ko) {
  Fr<R> ks = sph(k);
  R ct = phase_sample(s, rng.u64());
  R P[6];
  phase_eval(s, ct, P);
  R eq = -P[1] * I[1] / (P[0] * I[0]);
  R eu = -P[1] * I[2] / (P[0] * I[0]);
  R ee = sqrt(eq * eq + eu * eu);
  R phi;
  if (ee > 1e-12) {
    R gam = acos(clamp1(eq / ee));
    if (eu < 0) gam = M<R>::two_pi - gam;
    R ma = M<R>::four_pi * rng.u64() + gam - ee * sin(gam);
    R ea = kepler(ma, ee);
    phi = 0.5 * (ea - gam);
  } else {
    phi = M<R>::two_pi * rng.u64();
  }
  R st = sqrt0(1.0 - ct * ct);
  ko = ray_to_norm(V3<R>{st * cos(phi), st * sin(phi), ct}, ks);
  R Ip[4];
  update_I(s, I, ko, ks, Ip);
  R r = I[0] / Ip[0];
  for (int q = 0; q < 4; ++q) Io[q] = Ip[q] * r;
}

__global__ __launch_bounds__(256) void plain_kernel(const PlainParams* __restrict__ pp) {
  const PlainParams& p = *pp;
  const int64_t nid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)blockDim.x * gridDim.x;
  if (nid >= p.threads) return;
  Ctx c;
  c.p = &p;
  c.kp = p.kpath + 3 * p.maxscat * nid;
  c.xp = p.xpath + 3 * p.maxscat * nid;
  c.ip = p.idxpath + p.maxscat * nid;
  const GeoDev& g = p.geo;
  const int Nt = p.bins.Nt;
  int ub, ue, lb, le;
  if (g.kind == 0) { ub = Nt / 2; ue = Nt; lb = 0; le = Nt / 2; }
  else { ub = 0; ue = Nt; lb = 0; le = 0; }
  const bool can_dt = (g.kind == 1) || (g.nl == 1 && !g.isinf[0]);
  unsigned long long cnt[kCntN] = {};
  for (int64_t ray = nid; ray < p.N; ray += stride) {
    Rng rng{p.seed, p.ray0 + (uint64_t)ray, p.instance, 0};
    R I[4] = {p.I0[0], p.I0[1], p.I0[2], p.I0[3]};
    V3<R> k = {p.k0[0], p.k0[1], p.k0[2]};
    V3<R> x;
    if (g.kind == 0) {
      x = {0.0, 0.0, 0.0};
    } else {  // initial_x (sphere.jl:43-47)
      R r = g.tau[0] * sqrt(rng.u64());
      R ph = M<R>::two_pi * rng.u64();
      x = {r * cos(ph), r * sin(ph), sqrt(g.tau[0] * g.tau[0] - r * r)};
    }
    int idx = 0;
    int64_t n = 0;
    if (can_dt) {  // direct transmission (gpu.jl:196-201)
      R tp = distance(g, x, k, idx);
      R dt = tp <= p.tau0 ? exp(-tp) : 0.0;
      aadd(p.A + 0, I[0] * dt);
      for (int q = 0; q < 4; ++q) I[q] = I[q] * (1.0 - dt);
    }
    {  // initial_propagation (plane.jl:77-95, sphere.jl:54-67)
      R t;
      if (g.kind == 0) {
        if (g.isinf[0]) {
          t = -log(rng.u64());
        } else {
          R tp = distance(g, x, k, 0);
          if (can_dt && tp <= p.tau0) {
            t = -log(1.0 - rng.u64() * (1.0 - exp(-tp)));
          } else {
            t = tp + 1.0;
            for (int it = 0; t > tp && it < 100000000; ++it) t = -log(rng.u64());
          }
        }
      } else {
        R tp = 2.0 * x.z;
        if (tp <= p.tau0) {
          t = -log(1.0 - rng.u64() * (1.0 - exp(-tp)));
        } else {
          t = tp + 1.0;
          for (int it = 0; t > tp && it < 100000000; ++it) t = -log(1.0 - rng.u64());
        }
      }
      x = {x.x + t * k.x, x.y + t * k.y, x.z + t * k.z};
    }
    while (I[0] >= p.min_int && n < p.maxscat) {
      n += 1;
      st3(c.kp, n - 1, k);
      st3(c.xp, n - 1, x);
      c.ip[n - 1] = (int8_t)idx;
      R Ip[4];
      V3<R> kp, xp;
      int idxp = idx;
      for (long it = 0;; ++it) {
        scatter(p.sca[g.sca[idx]], I, k, rng, Ip, kp);
        R t = -log(rng.u64());
        xp = {x.x + t * kp.x, x.y + t * kp.y, x.z + t * kp.z};
        int pb;
        if (g.kind == 0) pb = xp.z > 0 ? -1 : ((!g.isinf[idx] && xp.z < -g.tau[idx]) ? 1 : 0);
        else pb = sqrt(dot(xp, xp)) > g.tau[0] ? -2 : 0;
        idxp = idx + pb;
        bool ok;
        if (pb == -2) ok = false;
        else if (pb == -1) {
          ok = idxp >= 0;
          if (ok) xp = {x.x, x.y, -g.tau[idx]};
        } else if (pb == 1) {
          ok = idxp <= g.nl - 1;
          if (ok) xp = {x.x, x.y, 0.0};
        } else ok = true;
        if (ok) break;
        ++cnt[kCntRej];
        if (it > 100000000) { idxp = idx; xp = x; break; }
      }
      R om = g.omega[idx];
      aadd(p.A + 1, (1.0 - om) * I[0]);
      for (int q = 0; q < 4; ++q) I[q] = I[q] * om;
      R Iref = 0.0, Itrans = 0.0;
      R dfirst, dlast;
      if (g.kind == 0) {
        dfirst = idx == 0 ? fabs(x.z) : __builtin_inf();
        dlast = (idx == g.nl - 1 && !g.isinf[idx]) ? g.tau[idx] - fabs(x.z) : __builtin_inf();
      } else {
        dfirst = g.tau[0] - sqrt(dot(x, x));
        dlast = __builtin_inf();
      }
      bool near_upper = dfirst <= p.tau0, near_lower = dlast <= p.tau0;
      if (p.compute_cb && (near_upper || near_lower)) collect_cb(c, I, k, x, idx, n, cnt[kCntCb], cnt[kCntCb1]);
      if (near_upper) Iref = collect_rt(c, I, k, x, idx, ub, ue, cnt[kCntRt]);
      if (near_lower) Itrans = collect_rt(c, I, k, x, idx, lb, le, cnt[kCntRt]);
      aadd(p.A + 2, Iref);
      aadd(p.A + 3, Itrans);
      R Iesc = Iref + Itrans;
      R s = I[0] - Iesc;
      for (int q = 0; q < 4; ++q) I[q] = Ip[q] * s / Ip[0];
      k = kp; x = xp; idx = idxp;
      ++cnt[kCntEvents];
    }
    if (!isfinite(I[0])) ++cnt[kCntNonfinite];
    aadd(p.A + 4, I[0]);
    ++cnt[kCntRays];
  }
  for (int q = 0; q < kCntN; ++q)
    if (cnt[q]) atomicAdd(p.counters + q, cnt[q]);
}

// Known-answer evaluation (mcrt_plan_eval) of the functions plain_kernel
// calls, with layer 0's scatterer.
__global__ void plain_eval_kernel(const PlainParams* __restrict__ pp, int what, int64_t n,
                                  const double* __restrict__ in, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PlainParams& p = *pp;
  const ScaView<R>& sca = p.sca[p.geo.sca[0]];
  const double* a = in + 16 * i;
  double* o = out + 4 * i;
  R r[4] = {0.0, 0.0, 0.0, 0.0};
  const Fr<R> f = {a[0], a[1], a[2], a[3]};
  switch (what) {
    case MCRT_EVAL_UPDATE_I: {
      const R I[4] = {a[4], a[5], a[6], a[7]};
      update_I(sca, I, V3<R>{a[8], a[9], a[10]}, f, r);
      break;
    }
    case MCRT_EVAL_UPDATE_E:
    case MCRT_EVAL_UPDATE_E_PATH: {
      Cx<R> E1 = {a[4], a[5]}, E2 = {a[6], a[7]};
      update_E(sca, E1, E2, V3<R>{a[8], a[9], a[10]}, f);
      r[0] = E1.re; r[1] = E1.im; r[2] = E2.re; r[3] = E2.im;
      break;
    }
    case MCRT_EVAL_DISTANCE:
      r[0] = distance(p.geo, V3<R>{a[0], a[1], a[2]}, V3<R>{a[3], a[4], a[5]}, (int)a[6]);
      r[1] = r[0];
      break;
    case MCRT_EVAL_PHASE:
    case MCRT_EVAL_PHASE_TRANSPORT: {
      R P[6];
      phase_eval(sca, a[0], P);
      r[0] = P[0]; r[1] = P[1]; r[2] = P[3]; r[3] = P[4];
      break;
    }
    case MCRT_EVAL_PHASE_DIAG: {
      R P[6];
      phase_eval(sca, a[0], P);
      r[0] = P[2]; r[1] = P[5]; r[2] = P[2]; r[3] = P[5];
      break;
    }
    case MCRT_EVAL_AMPLITUDE: {
      Cx<R> S11, S22;
      amplitude(sca, a[0], S11, S22);
      r[0] = S11.re; r[1] = S11.im; r[2] = S22.re; r[3] = S22.im;
      break;
    }
    case MCRT_EVAL_SAMPLE:
      r[0] = phase_sample(sca, a[0]);
      break;
    case MCRT_EVAL_KEPLER:
      r[0] = kepler(a[0], a[1]);
      break;
    default:
      break;
  }
  for (int c = 0; c < 4; ++c) o[c] = r[c];
}

}  // namespace

hipError_t launch_plain_eval(const PlainParams* p_dev, int what, int64_t n, const double* din,
                             double* dout, hipStream_t stream) {
  const int th = 256;
  hipLaunchKernelGGL(plain_eval_kernel, dim3((unsigned)((n + th - 1) / th)), dim3(th), 0, stream,
                     p_dev, what, n, din, dout);
  return hipGetLastError();
}

hipError_t launch_plain(const PlainParams* p_dev, int blocks, int threads, hipStream_t stream) {
  hipLaunchKernelGGL(plain_kernel, dim3(blocks), dim3(threads), 0, stream, p_dev);
  return hipGetLastError();
}

}  // namespace mcrt
