"""Pins of the coherent-backscattering half of the oracle against what the
mathematics fixes, computed without the oracle's frame algebra.

* E_from_I / I_from_E (src/utils.jl:119-163): fixed values, the V sign included.
* update_E, forward_E, backward_E (utils.jl:170-241): fields built in 3-D.  A
  Rayleigh dipole scatters a field E into E_s = i·(E − (E·k)k) (amplitude
  matrix diag(i, iμ), rayleigh.jl:20-23); a Mie sphere into
  S₁(E·n)n + S₂(E·p)p′ with n ⟂ the scattering plane, p = k × n, p′ = k′ × n.
  Components are taken in the basis (−ϕ̂, θ̂) of each direction (E₁ along
  −ϕ̂, E₂ along θ̂), phases included, at 1e-12.
* Reciprocity at exact backscattering: for transverse a, b the forward and
  reversed path operators satisfy b·T_a·a = a·T_b·b.
* collect_cb's n = 1 body (cpu.jl:131-158): the closed-form single-scattering
  Stokes vector of a dipole, in the bin's meridian basis, with the rotation as
  the reference writes it (see test_cb_single_scattering_closed_form).
* collect_cb's n > 1 body (cpu.jl:52-129): averaged over the interference
  phase, its I component is the ladder term update_I(I, k_f)₁·e^{−τ_a}.
* The same n = 1 bins from the transport (maximum_scattering_times = 1) against
  the closed-form single-scattering reflection, per (θb, ϕb).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import mcrt
import oracle
from statcheck import compare

rng = np.random.default_rng(20240611)
TOL = 1e-12


# ---------------------------------------------------------------- 3-D helpers
def frame(t, p):
    return np.array([math.cos(t), math.sin(t), math.cos(p), math.sin(p)])


def sph(k):
    """(cosθ, sinθ, cosϕ, sinϕ) of a unit vector, with the pole taken as ϕ = 0."""
    s = math.hypot(k[0], k[1])
    if s * s < 1e-24:
        return np.array([math.copysign(1.0, k[2]), 0.0, 1.0, 0.0])
    return np.array([k[2], math.sqrt(max(0.0, 1 - k[2] ** 2)), k[0] / s, k[1] / s])


def axes(f):
    ct, st, cp, sp = f
    th = np.array([ct * cp, ct * sp, -st])
    ph = np.array([-sp, cp, 0.0])
    k = np.array([st * cp, st * sp, ct])
    return th, ph, k


def to3(E, f):
    th, ph, _ = axes(f)
    return E[0] * (-ph) + E[1] * th


def from3(F, f):
    th, ph, _ = axes(f)
    return np.array([F @ (-ph), F @ th])


def dipole(F, kout):
    return 1j * (F - (F @ kout) * kout)


def mie_step(sca, F, kin, kout):
    n = np.cross(kin, kout)
    n /= np.linalg.norm(n)
    S1, S2 = oracle.amplitude(sca, float(kin @ kout))
    return S1 * (F @ n) * n + S2 * (F @ np.cross(kin, n)) * np.cross(kout, n)


def unit():
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def cvec():
    return rng.normal(size=2) + 1j * rng.normal(size=2)


def path(k0, n):
    """k0 then n-1 random directions, none near a pole or near (anti)parallel to
    its predecessor (where the scattering plane is undefined)."""
    ks = [k0]
    while len(ks) < n:
        k = unit()
        if abs(k[2]) < 0.98 and abs(k @ ks[-1]) < 0.95:
            ks.append(k)
    return ks


def oblique_frame():
    return frame(rng.uniform(0.6 * math.pi, 0.95 * math.pi), rng.uniform(0, 2 * math.pi))


RAY = oracle.Sca(mcrt.RayleighScatterer(1.0, 1.0))


def _mie():
    return oracle.Sca(mcrt.workloads.c2_mie(1801))


def _step(kind, sca):
    if kind == "rayleigh":
        return lambda F, kin, kout: dipole(F, kout)
    return lambda F, kin, kout: mie_step(sca, F, kin, kout)


# ------------------------------------------------------ (d) fixed field values
def test_I_from_E_fixed_values():
    """Q = |E₂|² − |E₁|², U = 2 Re E₂E₁*, V = −2 Im E₂E₁* (utils.jl:157-163)."""
    r = 1 / math.sqrt(2)
    cases = [((1, 0), (1, -1, 0, 0)), ((0, 1), (1, 1, 0, 0)), ((r, r), (1, 0, 1, 0)),
             ((r, -r), (1, 0, -1, 0)), ((r, 1j * r), (1, 0, 0, -1)), ((r, -1j * r), (1, 0, 0, 1)),
             ((2j, 0), (4, -4, 0, 0)), ((1, 1j), (2, 0, 0, -2))]
    for E, want in cases:
        got = oracle.I_from_E(np.array(E, dtype=complex))
        assert np.allclose(got, want, rtol=0, atol=1e-15), (E, got, want)


def test_E_from_I_fixed_values():
    """E₁ real and non-negative (ϕ₁ = 0); ϕ₂ = acos(U/√(I²−Q²)), mirrored to
    2π − ϕ₂ when V > 0 (utils.jl:119-150).  Unpolarised input gives the field of
    V = −1 (the reference's quirk)."""
    r = 1 / math.sqrt(2)
    cases = [((1, -1, 0, 0), (1, 0)), ((1, 1, 0, 0), (0, 1)), ((1, 0, 1, 0), (r, r)),
             ((1, 0, -1, 0), (r, -r)), ((1, 0, 0, 1), (r, -1j * r)), ((1, 0, 0, -1), (r, 1j * r)),
             ((1, 0, 0, 0), (r, 1j * r)), ((4, -4, 0, 0), (2, 0)), ((4, 4, 0, 0), (0, 2))]
    for I, want in cases:
        got = oracle.E_from_I(np.array(I, dtype=float))
        assert np.allclose(got, np.array(want, dtype=complex), rtol=0, atol=1e-15), (I, got, want)
    # V of the unpolarised input as written
    assert oracle.I_from_E(oracle.E_from_I([1.0, 0, 0, 0]))[3] == pytest.approx(-1.0, abs=1e-15)


# --------------------------------------------------- (b) fields against 3-D
@pytest.mark.parametrize("kind", ["rayleigh", "mie"])
def test_update_E_matches_3d_scattering(kind):
    sca = RAY if kind == "rayleigh" else _mie()
    step = _step(kind, sca)
    for _ in range(200):
        f = oblique_frame() if rng.uniform() < 0.5 else frame(rng.uniform(0.05, math.pi - 0.05),
                                                               rng.uniform(0, 2 * math.pi))
        kin = axes(f)[2]
        kout = path(kin, 2)[1]
        E = cvec()
        got, _, _ = oracle.update_E(sca, E, kout, f)
        want = from3(step(to3(E, f), kin, kout), sph(kout))
        assert np.max(np.abs(got - want)) < TOL * np.linalg.norm(E) * max(1, np.abs(want).max()), \
            (kind, got, want)


@pytest.mark.parametrize("kind", ["rayleigh", "mie"])
@pytest.mark.parametrize("n", [2, 3, 5])
def test_forward_and_backward_fields_match_3d_paths(kind, n):
    """forward_E walks k₂…k_n from the incident frame; backward_E walks
    −k_n…−k₂ (utils.jl:199-241).  Both renormalise to |E_i| and return
    enorm = −log(gain)."""
    sca = RAY if kind == "rayleigh" else _mie()
    step = _step(kind, sca)
    for _ in range(40):
        f0 = oblique_frame()
        k0 = axes(f0)[2]
        kp = path(k0, n)
        Ei = cvec()
        n0 = np.linalg.norm(Ei)
        for backward in (False, True):
            got, en = oracle.forward_E(sca, Ei, kp, f0, backward=backward)
            F = to3(Ei, f0)
            kin = k0
            seq = [-k for k in kp[:0:-1]] if backward else kp[1:]
            for kout in seq:
                F = step(F, kin, kout)
                kin = kout
            g = np.sqrt(np.sum(np.abs(F) ** 2))
            want = from3(F, sph(seq[-1])) * (n0 / g)
            assert np.max(np.abs(got - want)) < 1e-11 * n0, (kind, n, backward, got, want)
            assert abs(en + math.log(g / n0)) < 1e-11, (kind, n, backward)


@pytest.mark.parametrize("kind", ["rayleigh", "mie"])
def test_reciprocity_at_exact_backscatter(kind):
    """Reversing a path transposes its scattering operator: for transverse
    inputs a, b, b·(T_a a) = a·(T_b b) at k_f = −k₀, where T_a is forward_E
    followed by update_E into −k₀ from the last direction and T_b is
    backward_E followed by update_E into −k₀ from −k₂ (cpu.jl:65-76)."""
    sca = RAY if kind == "rayleigh" else _mie()
    for n in (2, 3, 4, 6):
        for _ in range(20):
            f0 = oblique_frame()
            k0 = axes(f0)[2]
            kp = path(k0, n)
            kf0 = -k0
            fo = sph(kf0)

            def tmap(E, backward):
                Eo, en = oracle.forward_E(sca, E, kp, f0, backward=backward)
                fl = sph(-kp[1]) if backward else sph(kp[-1])
                Ef, _, _ = oracle.update_E(sca, Eo, kf0, fl)
                return to3(Ef, fo) * math.exp(-en)  # the unnormalised operator's output

            a, b = cvec(), cvec()
            a3, b3 = to3(a, f0), to3(b, f0)
            lhs = b3 @ tmap(a, False)
            rhs = a3 @ tmap(b, True)
            scale = np.linalg.norm(a) * np.linalg.norm(b) * np.abs(tmap(a, False)).max()
            assert abs(lhs - rhs) < 1e-11 * max(scale, 1e-300), (kind, n, lhs, rhs)


# ----------------------------------------- (a) collect_cb n = 1, closed form
def _stokes(E1, E2):
    c = E2 * np.conj(E1)
    return np.array([abs(E1) ** 2 + abs(E2) ** 2, abs(E2) ** 2 - abs(E1) ** 2, 2 * c.real, -2 * c.imag])


def _field_of(q, u, v):
    """A unit field whose Stokes vector (in the convention above) is (1, q, u, v)."""
    E1 = math.sqrt(max(0.0, (1 - q) / 2))
    d = math.atan2(-v, u) if (u or v) else 0.0
    return np.array([E1, math.sqrt(max(0.0, (1 + q) / 2)) * np.exp(1j * d)])


def _slab(tau=0.7, omega=0.9, thetab_deg=(0.0, 1.0, 10.0, 33.0, 60.0, 85.0), Nphib=8, theta=math.pi,
          phi=0.0, kl=1.0):
    sca = mcrt.RayleighScatterer(omega, kl)
    g = mcrt.PlaneGeometry(theta, phi, [mcrt.Layer(d=tau * kl, scatterer=sca, tau=tau, omega=omega)])
    cfg = mcrt.Config(g, number_of_rays=1, compute_cb=True, Ntheta=8, Nphi=8,
                      thetab=np.deg2rad(np.asarray(thetab_deg, dtype=float)), Nphib=Nphib)
    return cfg, g


def test_cb_single_scattering_closed_form():
    """n = 1 at normal incidence: a dipole at depth z scatters the field E into
    √1.5·(E − (E·k_f)k_f) (P₁₁ normalised as rayleigh.jl:13), attenuated by
    e^{−z/cosθb}.  Its Stokes vector in the meridian basis of the bin,
    h = (sinϕb, −cosϕb, 0) and θ̂ = (cosθb cosϕb, cosθb sinϕb, −sinθb), is what
    collect_cb adds at θb = 0.  At θb > 0 the reference's final rotation
    (cpu.jl:149-153) uses get_basis's Eᵥ, whose second component is
    −k₃Eₕ₁ (utils.jl:60) rather than θ̂'s +k₃Eₕ₁: it gives cosψ = 1 and
    sinψ = −cosθb·sin2ϕb, so Q, U are mixed by (c² − s², 2cs) — a proper rotation
    only where sin2ϕb = 0.  That quirk is kept and pinned here."""
    cfg, g = _slab()
    S = oracle.Solve(cfg, g)
    k0 = oracle.ray_to_norm([0, 0, 1], frame(math.pi, 0.0))
    fin = sph(k0)
    for q, u, v in [(-1, 0, 0), (1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0.6, 0.0, 0.8),
                    (0.3, -0.8, 0.52)]:
        nrm = math.sqrt(q * q + u * u + v * v)
        q, u, v = q / nrm, u / nrm, v / nrm
        w = 0.37
        I = w * np.array([1.0, q, u, v])
        x = np.array([0.3, -0.2, -0.41])
        Icb = S.collect_cb_event(I, I, k0, x, 0, [k0], [x])
        F = to3(_field_of(q, u, v), fin) * math.sqrt(w)
        for j1, tb in enumerate(cfg.thetab):
            for j2, pb in enumerate(cfg.phib):
                kf = np.array([math.sin(tb) * math.cos(pb), math.sin(tb) * math.sin(pb), math.cos(tb)])
                Fs = math.sqrt(1.5) * (F - (F @ kf) * kf)
                h = np.array([math.sin(pb), -math.cos(pb), 0.0])
                t = np.array([math.cos(tb) * math.cos(pb), math.cos(tb) * math.sin(pb), -math.sin(tb)])
                St = _stokes(Fs @ h, Fs @ t) * math.exp(-0.41 / math.cos(tb))
                if tb > 0:
                    s = -math.cos(tb) * math.sin(2 * pb)
                    c2, s2 = 1 - s * s, 2 * s
                    St = np.array([St[0], c2 * St[1] + s2 * St[2], -s2 * St[1] + c2 * St[2], St[3]])
                got = Icb[:, j1, j2]
                assert np.max(np.abs(got - St)) < TOL * max(1.0, St[0]), (q, u, v, tb, pb, got, St)


def test_cb_single_scattering_transport_limit():
    """maximum_scattering_times = 1 on a thin Rayleigh slab at normal
    incidence, unpolarised light (mean of the ±Q runs): per (θb, ϕb), with
    μ = cosθb,
      ⟨Icb_I⟩ = ω·¾(1+μ²)·μ(1 − e^{−τ(1+1/μ)})/(1+μ),
      ⟨Icb_Q⟩ = −ω·¾(1−μ²)·(1 − s²)·μ(…)/(1+μ),  ⟨Icb_U⟩ = 2s/(1 − s²)·⟨Icb_Q⟩,
    s = −μ·sin2ϕb (the rotation of cpu.jl:149-153 as written); within 4.5σ
    per bin from 32 batch means, plus χ²/dof."""
    tau, omega = 0.5, 0.8
    cfg, g = _slab(tau, omega, thetab_deg=(0.0, 5.0, 20.0, 40.0, 60.0, 75.0), Nphib=8)
    cfg = cfg.replace(number_of_rays=6000, maximum_scattering_times=1)
    S = oracle.Solve(cfg, g)
    ra = S.simulate([1, -1, 0, 0], 11, nbatch=32)
    rb = S.simulate([1, 1, 0, 0], 12, nbatch=32)
    nrt = ra["nrt"]

    def per_batch(r):
        return r["batch_means"][:, 5 + nrt:].reshape(-1, cfg.Nphib, cfg.Nthetab, 4)

    u = 0.5 * (per_batch(ra) + per_batch(rb))  # (B, Nϕb, Nθb, 4)
    mean = u.mean(axis=0)
    sig = u.std(axis=0, ddof=1) / math.sqrt(u.shape[0])
    mu = np.cos(cfg.thetab)[None, :]
    pb = cfg.phib[:, None]
    depth = mu * (1 - np.exp(-tau * (1 + 1 / mu))) / (1 + mu)
    I = omega * 0.75 * (1 + mu**2) * depth * np.ones_like(pb)
    s = np.where(mu < 1.0, -mu * np.sin(2 * pb), 0.0)  # θb = 0: no quirk
    Q0 = -omega * 0.75 * (1 - mu**2) * depth
    Q = (1 - s * s) * Q0
    U = -2 * s * Q0
    for c, want in ((0, I), (1, Q), (2, U)):
        m, sg = mean[..., c], sig[..., c]
        live = sg > 0
        r = compare(m[live], want[live], sg[live], label=f"cb-single[{c}]")
        assert r["ok"] and r["max_abs_z"] < 4.5, r
        # where the batches carry no variance the value is exact (θb = 0: Q = U = 0)
        assert np.allclose(m[~live], want[~live], atol=1e-12)
    assert np.allclose(mean[..., 3], 0, atol=1e-14)


# ----------------------------------------- (c) collect_cb n > 1, ladder limit
@pytest.mark.parametrize("n", [2, 3, 5])
def test_cb_incoherent_average_is_ladder(n):
    """For one CB bin, four values of kl that move the interference phase
    ΔΦ = kl·(k_i + k_f)·(x_a − x_b) (cpu.jl:94) by quarter turns average the
    interference term out exactly, and the I component of
    I_from_E(E_a e^{iΔΦ} + E_b)·I_a/(I_a + I_b) (cpu.jl:110-114) becomes
    I_a = |E_af|², which for a fully polarised ray equals the ladder term
    update_I(I, k_f)₁·e^{−τ_a}: the field path (update_E) and the Stokes path
    (update_I) must agree in intensity at every k_f, not only at −k₀ where raa
    calibrates them (cpu.jl:68-73)."""
    for trial in range(12):
        th0 = math.pi if trial % 3 == 0 else rng.uniform(0.6 * math.pi, 0.95 * math.pi)
        f0 = frame(th0, 0.0 if trial % 3 == 0 else rng.uniform(0, 2 * math.pi))
        k0 = oracle.ray_to_norm([0, 0, 1], f0)
        kp = path(k0, n)
        tau = 3.0
        xs = [np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), -rng.uniform(0.2, 2.8)]) for _ in range(n)]
        tb, pb = math.radians(rng.uniform(2, 50)), rng.uniform(0, 2 * math.pi)
        I0 = np.array([1.0, -1.0, 0.0, 0.0]) if trial % 2 else np.array([1.0, 0.0, 0.6, 0.8])
        # the ray's Stokes vector at event n: that of its forward field, times a weight
        Ea, _ = oracle.forward_E(RAY, oracle.E_from_I(I0), kp, f0)
        I = 0.3 * oracle.I_from_E(Ea)
        kfs = sph(-k0)
        kf = oracle.ray_to_norm([math.sin(tb) * math.cos(pb), math.sin(tb) * math.sin(pb), math.cos(tb)],
                                kfs)
        d = (kp[0] + kf) @ (xs[-1] - xs[0])
        vals = []
        for j in range(4):
            kl = 7.0 + j * (math.pi / 2) / d
            cfg, g = _slab(tau, 0.9, thetab_deg=(math.degrees(tb),), Nphib=1, theta=math.acos(f0[0]),
                           phi=math.atan2(f0[3], f0[2]) % (2 * math.pi), kl=kl)
            cfg = cfg.replace(phib=np.array([pb]))
            S = oracle.Solve(cfg, g)
            vals.append(S.collect_cb_event(I0, I, kp[-1], xs[-1], 0, kp, xs)[:, 0, 0])
        ladder = oracle.update_I(RAY, I, kf, sph(kp[-1]))[0][0] * math.exp(-S.distance(xs[-1], kf))
        got = float(np.mean([v[0] for v in vals]))
        assert abs(got - ladder) < 1e-11 * abs(ladder), (n, trial, got, ladder)
        # and the interference itself is present: the four phases differ
        assert np.ptp([v[0] for v in vals]) > 1e-6 * ladder
