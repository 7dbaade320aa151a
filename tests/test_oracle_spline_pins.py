"""Pins of the oracle's SplineScatterer (src/scatterers/spline.jl) against what
the mathematics fixes, independent of the oracle's own formulas.

The reference's lookups read `self.NSPL`, a name defined nowhere, so they throw
on first use; NSPL is read as Nθ - 1 (the table's last node) and everything
else is kept as written: the phase matrix (spline.jl:33-57) interpolates with
weights (msn+1-sn, sn-msn) on nodes msn, msn+1 (1-based, msn = trunc(sn) + 1),
which extrapolates one node step: for a table linear in θ the lookup returns
the line one node step back, F(θ - Δθ), and at θ = π twice the last node.  The
amplitude lookup (spline.jl:60-76) is Mie's (mie.jl:76-92).

 * lookup: a table linear in θ gives F(θ - Δθ) exactly, 2·F[end] at θ = π;
 * element placement: with constant tables the rotations into and out of the
   scattering plane drop out of frame-free invariants of update_I
   (utils.jl:91-112): out₀ = P11 and |(Q, U)| = |P12| for unpolarised light,
   out₃ = P44 and |(Q, U)| = |P34| for circular light, and the Q and U inputs
   (rotated by the same 2ϕ) give Σ|(Q, U)|² = P22² when P33 = 0;
 * amplitude: bitwise the Mie lookup on the same S tables;
 * sampling: the csrn inversion of F11 (spline.jl:19-31, as Mie's);
 * transport: energy telescoping ΣA = 1 per ray on a Spline sphere and a
   Rayleigh-over-Spline slab.
"""
from __future__ import annotations

import math

import numpy as np

import mcrt
import oracle
from mcrt.scatterers import MieScatterer, SplineScatterer

rng = np.random.default_rng(2027)


def _spline(tables, S11=None, S22=None, rp=200):
    N = len(tables[0])
    S11 = np.ones(N, complex) if S11 is None else S11
    S22 = np.ones(N, complex) if S22 is None else S22
    F11 = tables[0] if np.all(tables[0] > 0) else np.ones(N)  # csrn needs F11·sinθ ≥ 0
    csrn = mcrt.calculate_csrn(F11, rp)
    return SplineScatterer(0.9, 1.2, csrn, *tables, S11, S22)


def _unit():
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def _frame():
    t, p = rng.uniform(0.05, math.pi - 0.05), rng.uniform(0, 2 * math.pi)
    return np.array([math.cos(t), math.sin(t), math.cos(p), math.sin(p)])


def test_spline_lookup_of_a_linear_table_is_one_node_step_back():
    N = 37
    th = np.linspace(0.0, math.pi, N)
    dth = math.pi / (N - 1)
    a = rng.uniform(0.5, 2.0, 6)
    b = rng.uniform(-0.3, 0.3, 6)
    tab = [a[c] + b[c] * th for c in range(6)]
    sca = oracle.Sca(_spline(tab))
    for t in np.concatenate([rng.uniform(0.0, math.pi * (1 - 1e-9), 400), th[:-1] + 1e-7]):
        P = oracle.phase_eval(sca, math.cos(t))
        tt = math.acos(math.cos(t))  # the θ the lookup sees
        want = a + b * (tt - dth)
        assert np.allclose(P, want, rtol=0, atol=1e-12 * np.abs(a).max()), (t, P, want)
    # θ = π: sn = NSPL, msn + 1 exceeds the table, only (msn+1-sn)·F[end] = 2·F[end]
    P = oracle.phase_eval(sca, -1.0)
    assert np.allclose(P, 2 * np.array([tab[c][-1] for c in range(6)]), rtol=1e-14)


def test_spline_phase_elements_are_placed_as_written():
    """[P11 P12 0 0; P12 P22 0 0; 0 0 P33 P34; 0 0 -P34 P44] (spline.jl:52-55)."""
    N = 19
    for _ in range(40):
        v = rng.uniform(-1.0, 1.0, 6)
        v[0] = abs(v[0]) + 0.5
        tab = [np.full(N, x) for x in v]
        P11, P12, P22, P33, P34, P44 = v
        sca = oracle.Sca(_spline(tab))
        tab0 = list(tab)
        tab0[3] = np.zeros(N)  # P33 = 0 isolates P22
        sca0 = oracle.Sca(_spline(tab0))
        for _ in range(10):
            f = _frame()
            k = _unit()
            # constant tables: the weights sum to one away from θ = π
            kin = np.array([f[1] * f[2], f[1] * f[3], f[0]])
            if k @ kin < -0.999:
                continue
            o, _, _ = oracle.update_I(sca, [1.0, 0.0, 0.0, 0.0], k, f)
            assert abs(o[0] - P11) < 1e-12 and abs(math.hypot(o[1], o[2]) - abs(P12)) < 1e-12
            assert abs(o[3]) < 1e-15
            o, _, _ = oracle.update_I(sca, [0.0, 0.0, 0.0, 1.0], k, f)
            assert abs(o[3] - P44) < 1e-12 and abs(math.hypot(o[1], o[2]) - abs(P34)) < 1e-12
            assert abs(o[0]) < 1e-15
            oq, _, _ = oracle.update_I(sca0, [0.0, 1.0, 0.0, 0.0], k, f)
            ou, _, _ = oracle.update_I(sca0, [0.0, 0.0, 1.0, 0.0], k, f)
            assert abs(oq[1] ** 2 + oq[2] ** 2 + ou[1] ** 2 + ou[2] ** 2 - P22 ** 2) < 1e-12
            # the Q and U inputs also reach I through P12 and V through P34 (zero here)
            assert abs(oq[0] ** 2 + ou[0] ** 2 - P12 ** 2) < 1e-12
            o2, _, _ = oracle.update_I(sca, [0.0, 1.0, 0.0, 0.0], k, f)
            o3, _, _ = oracle.update_I(sca, [0.0, 0.0, 1.0, 0.0], k, f)
            assert abs(o2[3] ** 2 + o3[3] ** 2 - P34 ** 2) < 1e-12
            assert abs(sum(x[1] ** 2 + x[2] ** 2 for x in (o2, o3)) - P22 ** 2 - P33 ** 2) < 1e-12


def test_spline_amplitude_is_the_mie_lookup():
    m = mcrt.workloads.c2_mie(181)
    sp = mcrt.workloads.spline_from(m)
    a, b = oracle.Sca(m), oracle.Sca(sp)
    for mu in np.concatenate([rng.uniform(-1, 1, 300), [-1.0, 1.0]]):
        assert oracle.amplitude(a, mu) == oracle.amplitude(b, mu)
        assert oracle.phase_sample(a, (mu + 1) / 2.0001 + 1e-6) == \
            oracle.phase_sample(b, (mu + 1) / 2.0001 + 1e-6)


def test_spline_transport_energy_telescopes():
    m = mcrt.workloads.c2_mie(181)
    sp = mcrt.workloads.spline_from(m)
    g = mcrt.SphereGeometry(R=2.0 * sp.kl, scatterer=sp)
    cfg = mcrt.Config(g, number_of_rays=64, Ntheta=30, Nphi=8, compute_cb=True)
    S = oracle.Solve(cfg, g)
    for ray in range(8):
        A = S.trace_ray([1.0, -1.0, 0.0, 0.0], 5, ray)[0]
        assert abs(sum(A) - 1.0) < 1e-12, A
    ray_s = mcrt.RayleighScatterer(1.0, 1.0)
    g = mcrt.PlaneGeometry(math.pi, 0.0, [mcrt.Layer(d=0.5, scatterer=ray_s, tau=0.5, omega=1.0),
                                          mcrt.Layer(d=sp.kl, scatterer=sp, tau=1.0, omega=sp.omega)])
    cfg = mcrt.Config(g, number_of_rays=64, Ntheta=30, Nphi=8, compute_cb=True)
    S = oracle.Solve(cfg, g)
    for ray in range(8):
        A = S.trace_ray([1.0, 0.0, 1.0, 0.0], 6, ray)[0]
        assert abs(sum(A) - 1.0) < 1e-12, A
