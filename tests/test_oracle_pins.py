"""Pins of the oracle's deterministic parts (CPU only).

Each test checks a part of oracle/mcrt_oracle.c against something the
reference's formulas do not themselves supply: published known-answer
vectors, closed forms, independent library routines (numpy/scipy/mpmath),
physical laws, or brute force on tiny inputs.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import mcrt
import oracle

GOLDEN = Path(__file__).parent / "golden"
rng = np.random.default_rng(1234)


# ----------------------------------------------------------------- Philox
def test_philox_known_answers():
    """Random123 known-answer vectors of philox4x32-10 (kat_vectors)."""
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E,
                                                                0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                         [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniform_conversion():
    seed, inst = 0x0123456789ABCDEF, 3
    for ray, draw in [(0, 0), (5, 7), (2**40 + 3, 11)]:
        w = oracle.philox([ray & 0xFFFFFFFF, ray >> 32, draw, inst], [seed & 0xFFFFFFFF, seed >> 32])
        bits = (w[0] << 32) | w[1]
        assert oracle.uniform(seed, inst, ray, draw) == ((bits >> 12) + 0.5) * 2.0**-52
        assert oracle.uniform_f32(seed, inst, ray, draw) == np.float32(((bits >> 41) + 0.5) * 2.0**-23)
    u = np.array([oracle.uniform(42, 0, r, d) for r in range(200) for d in range(10)])
    assert u.min() > 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.02


# ------------------------------------------------------- Gauss–Legendre
@pytest.mark.parametrize("n", [1, 2, 5, 24, 90, 180])
def test_gausslegendre_vs_numpy(n):
    x, w = oracle.gausslegendre(n)
    xn, wn = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xn)) < 2e-15
    assert np.max(np.abs(w - wn)) < 2e-14 * max(1, wn.max())
    assert abs(w.sum() - 2) < 1e-13
    # exact for polynomials of degree ≤ 2n-1
    for deg in range(0, 2 * n, max(1, n // 3)):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.dot(w, x**deg) - exact) < 1e-13


def test_initial_weights_normalisation():
    for kind, Nt, Np in [(0, 180, 48), (1, 180, 48), (0, 64, 48), (1, 65, 8)]:
        ct, st, w = oracle.initial_weights(kind, Nt, Np)
        assert abs(w.sum() * Np - 1) < 1e-13  # ΔΩ/4π over the sphere
        assert np.all(np.diff(ct) > 0)
        assert np.allclose(st, np.sqrt(1 - ct**2), rtol=0, atol=0)
        if kind == 0:
            h = Nt // 2
            assert np.all(ct[:h] < 0) and np.all(ct[h:] > 0)
            assert abs(w[:h].sum() * Np - 0.5) < 1e-13


# ------------------------------------------------------------- frames
def _Rz(p):
    return np.array([[math.cos(p), -math.sin(p), 0], [math.sin(p), math.cos(p), 0], [0, 0, 1]])


def _Ry(t):
    return np.array([[math.cos(t), 0, math.sin(t)], [0, 1, 0], [-math.sin(t), 0, math.cos(t)]])


def _frame(t, p):
    return np.array([math.cos(t), math.sin(t), math.cos(p), math.sin(p)])


def _unit():
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def test_ray_to_norm_is_RzRy():
    for _ in range(50):
        t, p = rng.uniform(0, math.pi), rng.uniform(0, 2 * math.pi)
        x = rng.normal(size=3)
        R = _Rz(p) @ _Ry(t)
        assert np.allclose(oracle.ray_to_norm(x, _frame(t, p)), R @ x, atol=1e-14)
        assert np.allclose(oracle.norm_to_ray(x, _frame(t, p)), R.T @ x, atol=1e-14)
        back = oracle.norm_to_ray(oracle.ray_to_norm(x, _frame(t, p)), _frame(t, p))
        assert np.allclose(back, x, atol=1e-14)


def test_spherical_from_cartesian():
    for _ in range(50):
        k = _unit()
        ct, st, cp, sp = oracle.spherical_from_cartesian(k)
        th = math.acos(k[2])
        ph = math.atan2(k[1], k[0])
        assert np.allclose([ct, st, cp, sp], [math.cos(th), math.sin(th), math.cos(ph), math.sin(ph)],
                           atol=1e-14)
    # polar branch
    assert list(oracle.spherical_from_cartesian([1e-13, 0, -1])) == [-1.0, 0.0, 1.0, 0.0]


def test_get_basis_and_its_second_component():
    for _ in range(50):
        k = _unit()
        Eh, Ev = oracle.get_basis(k)
        assert abs(np.dot(Eh, k)) < 1e-14 and abs(np.linalg.norm(Eh) - 1) < 1e-14
        assert Eh[2] == 0.0
        # as written (utils.jl:60): Ev = (-k₃Eₕ₂, -k₃Eₕ₁, -√(1-k₃²)); the θ-hat
        # vector would have +k₃Eₕ₁ in the second slot
        assert np.allclose(Ev, [-k[2] * Eh[1], -k[2] * Eh[0], -math.sqrt(1 - k[2] ** 2)], atol=1e-15)
    k = np.array([0.6, 0.0, 0.8])  # Eₕ₁ = 0: Ev is then the true θ-hat
    Eh, Ev = oracle.get_basis(k)
    assert abs(np.dot(Ev, k)) < 1e-15


def test_rotate_invariants():
    for _ in range(20):
        I = rng.normal(size=4)
        a = rng.uniform(0, math.pi)
        o = oracle.rotate(I, math.cos(2 * a), math.sin(2 * a))
        assert o[0] == I[0] and o[3] == I[3]
        assert abs(o[1] ** 2 + o[2] ** 2 - I[1] ** 2 - I[2] ** 2) < 1e-13


# ---------------------------------------------------- scatterer sampling
def test_rayleigh_inverse_cdf():
    """μ(u) from Cardano's formula inverts F(μ) = 1/2 + 3/8(μ + μ³/3)."""
    sca = oracle.Sca(mcrt.RayleighScatterer(1.0, 1.0))
    for u in np.linspace(1e-6, 1 - 1e-6, 101):
        mu = oracle.phase_sample(sca, u)
        F = 0.5 + 3 / 8 * (mu + mu**3 / 3)
        assert abs(F - (1 - u)) < 1e-12 or abs(F - u) < 1e-12
    # orientation: small u → μ near +1 (rn = 2(1-2u) > 0)
    assert oracle.phase_sample(sca, 1e-9) > 0.999


def test_rayleigh_phase_matrix_normalised():
    sca = oracle.Sca(mcrt.RayleighScatterer(1.0, 1.0))
    mu, w = np.polynomial.legendre.leggauss(16)
    P = np.array([oracle.phase_eval(sca, m) for m in mu])
    assert abs(0.5 * np.dot(w, P[:, 0]) - 1.0) < 1e-14  # ∫P11 dΩ/4π = 1
    assert np.allclose(P[:, 1] / P[:, 0], -(1 - mu**2) / (1 + mu**2), atol=1e-15)


def _rayleigh_check(k_in_frame, k_out, I_in, e_in):
    """Dipole law: out intensity 1.5(1-(e·k)²), fully polarised along
    E_s = e - (e·k)k; Q'/I' = 1 - 2 (Ê_s·Eₕ)²."""
    sca = oracle.Sca(mcrt.RayleighScatterer(1.0, 1.0))
    out, Eh, Ev = oracle.update_I(sca, I_in, k_out, k_in_frame)
    ek = float(np.dot(e_in, k_out))
    assert abs(out[0] - 1.5 * (1 - ek**2)) < 1e-12
    dop = math.sqrt(out[1] ** 2 + out[2] ** 2 + out[3] ** 2) / out[0]
    assert abs(dop - 1) < 1e-9
    Es = e_in - ek * k_out
    Es /= np.linalg.norm(Es)
    f = float(np.dot(Es, Eh)) ** 2
    assert abs(out[1] / out[0] - (1 - 2 * f)) < 1e-9
    assert abs(abs(out[2]) / out[0] - 2 * math.sqrt(max(f * (1 - f), 0))) < 1e-7


def test_update_I_rayleigh_dipole_law():
    """update_I against dipole scattering in absolute 3-D space.  Incident Q=+1
    is polarised along θ-hat, Q=-1 along ϕ-hat (E_from_I puts Q>0 into E₂)."""
    signs = []
    for _ in range(200):
        t, p = rng.uniform(0.05, math.pi - 0.05), rng.uniform(0, 2 * math.pi)
        f = _frame(t, p)
        th_hat = oracle.ray_to_norm([1, 0, 0], f)
        ph_hat = oracle.ray_to_norm([0, 1, 0], f)
        k_out = _unit()
        _rayleigh_check(f, k_out, [1, 1, 0, 0], th_hat)
        _rayleigh_check(f, k_out, [1, -1, 0, 0], ph_hat)
        # U = +1 is polarised along one diagonal, the same one for every draw
        sca = oracle.Sca(mcrt.RayleighScatterer(1.0, 1.0))
        o, _, _ = oracle.update_I(sca, [1, 0, 1, 0], k_out, f)
        d1 = (th_hat + ph_hat) / math.sqrt(2)
        d2 = (th_hat - ph_hat) / math.sqrt(2)
        i1 = 1.5 * (1 - np.dot(d1, k_out) ** 2)
        i2 = 1.5 * (1 - np.dot(d2, k_out) ** 2)
        if abs(i1 - i2) > 1e-3:
            signs.append(1 if abs(o[0] - i1) < 1e-12 else (2 if abs(o[0] - i2) < 1e-12 else 0))
    assert 0 not in signs and len(set(signs)) == 1


def test_update_I_unpolarised_intensity_mie():
    """I'₁ for unpolarised light is F11 at the scattering angle, in any frame."""
    s = mcrt.workloads.c2_mie(1801)
    sca = oracle.Sca(s)
    for _ in range(50):
        t, p = rng.uniform(0.05, math.pi - 0.05), rng.uniform(0, 2 * math.pi)
        f = _frame(t, p)
        k_in = oracle.ray_to_norm([0, 0, 1], f)
        k_out = _unit()
        a, _, _ = oracle.update_I(sca, [1, 1, 0, 0], k_out, f)
        b, _, _ = oracle.update_I(sca, [1, -1, 0, 0], k_out, f)
        mu = float(np.dot(k_in, k_out))
        P = oracle.phase_eval(sca, mu)
        assert abs(0.5 * (a[0] + b[0]) - P[0]) < 1e-10 * max(1, P[0])


def test_mie_phase_lookup_and_amplitude_shift():
    s = mcrt.workloads.c2_mie(1801)
    sca = oracle.Sca(s)
    N = s.Ntheta
    for i in [0, 1, 100, 900, 1799]:
        th = i * math.pi / (N - 1)
        P = oracle.phase_eval(sca, math.cos(th))
        assert abs(P[0] - s.F11[i]) < 1e-9 * s.F11.max()
        # midpoint → average of the two nodes (correct linear interpolation)
        thm = (i + 0.5) * math.pi / (N - 1)
        Pm = oracle.phase_eval(sca, math.cos(thm))
        assert abs(Pm[0] - 0.5 * (s.F11[i] + s.F11[i + 1])) < 1e-8 * s.F11.max()
    # amplitude_matrix weights (msn+1-sn, sn-msn): one-step extrapolation
    S1, S2 = oracle.amplitude(sca, -1.0)
    assert abs(S1 - 2 * s.S11[-1]) < 1e-12 * abs(s.S11[-1]) * 2
    thm = 100.5 * math.pi / (N - 1)
    S1, _ = oracle.amplitude(sca, math.cos(thm))
    assert abs(S1 - (1.5 * s.S11[100] - 0.5 * s.S11[101])) < 1e-8 * abs(s.S11).max()


# ------------------------------------------------------------ fields
def test_E_I_round_trip():
    for _ in range(100):
        v = _unit()
        I = np.array([1.0, v[0], v[1], v[2]])
        E = oracle.E_from_I(I)
        assert np.allclose(oracle.I_from_E(E), I, atol=1e-12)
    # unpolarised input maps to a circular field (utils.jl:119-150 as written)
    Iu = oracle.I_from_E(oracle.E_from_I([1.0, 0, 0, 0]))
    assert abs(abs(Iu[3]) - 1) < 1e-12


def test_update_E_intensity_is_dipole_law():
    """|E'|² = 1 - (e·k)² for Rayleigh (|S11|² = 1, |S22|² = μ²)."""
    sca = oracle.Sca(mcrt.RayleighScatterer(1.0, 1.0))
    for _ in range(100):
        t, p = rng.uniform(0.05, math.pi - 0.05), rng.uniform(0, 2 * math.pi)
        f = _frame(t, p)
        th_hat = oracle.ray_to_norm([1, 0, 0], f)
        ph_hat = oracle.ray_to_norm([0, 1, 0], f)
        k_out = _unit()
        for E, e in (([0, 1], th_hat), ([1, 0], ph_hat)):
            Eo, _, _ = oracle.update_E(sca, np.array(E, complex), k_out, f)
            assert abs(np.sum(np.abs(Eo) ** 2) - (1 - np.dot(e, k_out) ** 2)) < 1e-12


def test_forward_backward_E_norms():
    """forward_E/backward_E renormalise each step: output norm = input norm and
    enorm = -log of the unnormalised gain (telescoping)."""
    s = mcrt.workloads.c2_mie(1801)
    sca = oracle.Sca(s)
    f0 = _frame(math.pi, 0.0)
    kp = [oracle.ray_to_norm([0, 0, 1], f0)] + [_unit() for _ in range(6)]
    Ei = oracle.E_from_I([1, -1, 0, 0])
    Ea, ea = oracle.forward_E(sca, Ei, kp, f0)
    Eb, eb = oracle.forward_E(sca, Ei, kp, f0, backward=True)
    assert abs(np.linalg.norm(Ea) - 1) < 1e-12 and abs(np.linalg.norm(Eb) - 1) < 1e-12
    # unnormalised product of single steps
    E = Ei.copy()
    fr = f0
    for k in kp[1:]:
        E, _, _ = oracle.update_E(sca, E, k, fr)
        fr = oracle.spherical_from_cartesian(k)
    assert abs(ea + math.log(np.linalg.norm(E))) < 1e-10
    assert np.allclose(Ea, E / np.linalg.norm(E), atol=1e-12)


# ------------------------------------------------------------- Kepler
def test_kepler_residual_and_mpmath():
    import mpmath as mp
    for _ in range(200):
        e = rng.uniform(0, 0.999)
        M = rng.uniform(0, 6 * math.pi)
        E = oracle.kepler(M, e)
        assert abs(E - e * math.sin(E) - M) < 1e-11
    for e, M in [(0.5, 1.0), (0.9, 4.0), (0.99, 0.1)]:
        ref = float(mp.findroot(lambda E: E - e * mp.sin(E) - M, M))
        assert abs(oracle.kepler(M, e) - ref) < 1e-11


# ------------------------------------------------------------ geometry
def test_distances_closed_form_and_bisection():
    g = mcrt.PlaneGeometry(math.pi, 0.0, [mcrt.Layer(d=2.0, scatterer=mcrt.RayleighScatterer(1.0, 1.0))])
    cfg = mcrt.Config(g, number_of_rays=1, Ntheta=8, Nphi=8)
    S = oracle.Solve(cfg, g)
    for _ in range(20):
        x = np.array([0.3, -0.2, -rng.uniform(0, 2)])
        k = _unit()
        d = S.distance(x, k)
        y = x + d * k
        if k[2] > 0:
            assert abs(y[2]) < 1e-12
        else:
            assert abs(y[2] + 2.0) < 1e-12
    sph = mcrt.SphereGeometry(R=3.0, scatterer=mcrt.RayleighScatterer(1.0, 1.0))
    S = oracle.Solve(mcrt.Config(sph, number_of_rays=1, Ntheta=8, Nphi=8), sph)
    for _ in range(20):
        x = _unit() * rng.uniform(0, 2.9)
        k = _unit()
        d = S.distance(x, k)
        lo, hi = 0.0, 10.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if np.linalg.norm(x + mid * k) < 3.0:
                lo = mid
            else:
                hi = mid
        assert abs(d - lo) < 1e-12


# ------------------------------------------------------------ tables
def test_bhmie_golden():
    G = json.loads((GOLDEN / "bhmie_golden.json").read_text())
    for c in G:
        r = oracle.bhmie(c["x"], complex(*c["m"]), c["Ntheta"])
        idx = [a["i"] for a in c["angles"]]
        for k in ("F11", "F12", "F33", "F34"):
            g = np.array([a[k] for a in c["angles"]])
            assert np.max(np.abs(g - r[k][idx])) <= 1e-12 * np.max(np.abs(g)), (c["x"], k)
        for k in ("S1", "S2"):
            g = np.array([complex(*a[k]) for a in c["angles"]])
            assert np.max(np.abs(g - r[k][idx])) <= 1e-12 * np.max(np.abs(g)), (c["x"], k)
        for k in ("q_ext", "q_sca", "q_back"):
            assert abs(r[k] - c[k]) <= 1e-12 * abs(c[k]), (c["x"], k)
        assert abs(r["asymm"] - c["asymm"]) < 1e-12


def test_bhmie_vs_scipy_bessel_series():
    """Independent algorithm: a_n, b_n from scipy spherical Bessel functions
    (Bohren & Huffman eq. 4.53), truncated at the same nstop."""
    from scipy.special import spherical_jn, spherical_yn
    for x, m in [(5.71198664289, 1.5 + 0.001j), (1.0, 1.33 + 0j), (12.0, 1.6 + 0.05j)]:
        r = oracle.bhmie(x, m, 181)
        nstop = int(math.floor(max(x + 4 * x ** (1 / 3) + 2, abs(m) * x)))
        n = np.arange(1, nstop + 1)
        mx = m * x
        jx, djx = spherical_jn(n, x), spherical_jn(n, x, derivative=True)
        yx, dyx = spherical_yn(n, x), spherical_yn(n, x, derivative=True)
        # ψ_n(z) = z j_n(z); ξ_n = x h_n^(1)(x)
        psi, dpsi = x * jx, jx + x * djx
        xi, dxi = x * (jx + 1j * yx), (jx + 1j * yx) + x * (djx + 1j * dyx)
        import mpmath as mp
        jm = np.array([complex(mp.sqrt(mp.pi / (2 * mx)) * mp.besselj(k + 0.5, mx)) for k in n])
        djm = np.array([complex(mp.sqrt(mp.pi / (2 * mx)) * mp.besselj(k + 0.5, mx, derivative=1)
                                - mp.sqrt(mp.pi / (2 * mx)) * mp.besselj(k + 0.5, mx) / (2 * mx))
                        for k in n])
        psim, dpsim = mx * jm, jm + mx * djm
        a = (m * psim * dpsi - psi * dpsim) / (m * psim * dxi - xi * dpsim)
        b = (psim * dpsi - m * psi * dpsim) / (psim * dxi - m * xi * dpsim)
        qext = 2 / x**2 * np.sum((2 * n + 1) * (a + b).real)
        qsca = 2 / x**2 * np.sum((2 * n + 1) * (abs(a) ** 2 + abs(b) ** 2))
        assert abs(r["q_ext"] - qext) < 1e-9 * qext
        assert abs(r["q_sca"] - qsca) < 1e-9 * qsca
    # real m: no absorption
    r = oracle.bhmie(3.0, 1.33 + 0j, 181)
    assert abs(r["q_abs"]) < 1e-13


def test_bhmie_rayleigh_limit():
    r = oracle.bhmie(0.01, 1.5 + 0j, 181)
    mu = np.cos(np.linspace(0, math.pi, 181))
    assert np.allclose(r["F12"] / r["F11"], -(1 - mu**2) / (1 + mu**2), atol=1e-5)
    assert np.allclose(r["F11"] / r["F11"][90], (1 + mu**2), atol=1e-4)


def test_csrn_properties():
    s = mcrt.workloads.c2_mie(1801)
    c = oracle.calculate_csrn(s.F11, 1000)
    assert np.all(np.diff(c) > 0)
    assert c[-1] == 1.0 and abs(c[0] + math.cos(math.pi / 1000)) < 1e-12
    # Rayleigh-shaped F11: csrn converges to the closed-form inverse CDF
    th = np.linspace(0, math.pi, 1801)
    F = 0.75 * (1 + np.cos(th) ** 2)
    cr = oracle.calculate_csrn(F, 1000)
    sca = oracle.Sca(mcrt.RayleighScatterer(1.0, 1.0))
    u = np.arange(1001) / 1000
    exact = np.array([oracle.phase_sample(sca, 1 - ui) if 0 < ui < 1 else (2 * ui - 1) for ui in u])
    assert np.max(np.abs(cr - exact)) < 5e-3


def test_csrn_independent_numpy_restatement():
    s = mcrt.workloads.c2_mie(1801)
    n = 1000
    N = len(s.F11)
    xv = np.arange(n + 1) / (n / math.pi)
    F = np.interp(xv, np.linspace(0, math.pi, N), s.F11)
    y = np.cumsum(F * np.sin(xv))
    y = 1 - y / y[-1]
    knots = y[::-1][1:]
    vals = (-np.cos(math.pi - xv))[::-1][1:]
    ref = np.interp(np.arange(n + 1) / n, knots, vals)
    assert np.max(np.abs(oracle.calculate_csrn(s.F11, n) - ref)) < 1e-12


# ------------------------------------------------------------ transport
def test_energy_telescopes_per_ray():
    for make in (lambda: mcrt.workloads.plane_rayleigh_tau1(1, compute_cb=True),
                 lambda: mcrt.workloads.sphere_mie_cb(1)):
        cfg, g, seed = make()
        S = oracle.Solve(cfg, g)
        for ray in range(5):
            A, _, _, cnt = S.trace_ray([1, -1, 0, 0], seed, ray)
            assert abs(A.sum() - 1.0) < 1e-12, A


def test_thread_count_independence():
    cfg, g, seed = mcrt.workloads.plane_rayleigh_tau1(64, compute_cb=True)
    S = oracle.Solve(cfg, g)
    a = S.simulate([1, -1, 0, 0], seed, nthreads=1, nbatch=8)
    b = S.simulate([1, -1, 0, 0], seed, nthreads=7, nbatch=8)
    assert np.array_equal(a["Irt"], b["Irt"]) and np.array_equal(a["Icb"], b["Icb"])
    assert np.array_equal(a["A"], b["A"])


def _single_scattering(tau, omega, N, grid):
    cfg, g, seed = mcrt.workloads.plane_rayleigh_tau1(N, tau=tau, omega=omega, max_orders=1,
                                                      gl_grid=grid == "gl", Ntheta=40, Nphi=8)
    return cfg, g, seed


@pytest.mark.parametrize("grid", ["gl", "mid"])
def test_rayleigh_single_scattering_limit(grid):
    """maximum_scattering_times = 1: the reflected intensity of unpolarised light
    at normal incidence is R₁ = ω·¾(1+μ²)·(1 - e^{-τ(1+1/μ)})/(4(1+μ)) in the
    reference's normalisation Irt/|4 cosθ·inorm|, and Q/I = -(1-μ²)/(1+μ²).
    Checked per polar bin within 4σ (σ from batch means) plus χ²/dof."""
    from statcheck import compare
    tau, omega = 0.7, 0.9
    cfg, g, seed = _single_scattering(tau, omega, 8000, grid)
    S = oracle.Solve(cfg, g)
    h = cfg.Ntheta // 2
    mu = cfg.costheta[h:]
    norm = 1.0 / np.abs(4 * cfg.costheta[h:] * cfg.inorm[h:])

    def per_batch(res):
        m = res["batch_means"][:, 5:5 + res["nrt"]].reshape(-1, cfg.Nphi, cfg.Ntheta, 4)
        return m[:, :, h:, :].mean(axis=1) * norm[None, :, None]  # (B, Nθ/2, 4)

    pa = per_batch(S.simulate([1, -1, 0, 0], seed, nbatch=32))
    pb = per_batch(S.simulate([1, 1, 0, 0], seed + 1, nbatch=32))
    u = 0.5 * (pa + pb)  # unpolarised = average of the ±Q columns
    R = u[:, :, 0].mean(axis=0)
    sig = u[:, :, 0].std(axis=0, ddof=1) / math.sqrt(u.shape[0])
    R1 = omega * 0.75 * (1 + mu**2) * (1 - np.exp(-tau * (1 + 1 / mu))) / (4 * (1 + mu))
    r = compare(R, R1, sig, label=f"R1/{grid}")
    assert r["ok"] and r["max_abs_z"] < 4.5, r
    Q = u[:, :, 1].mean(axis=0)
    assert np.allclose(Q / R, -(1 - mu**2) / (1 + mu**2), atol=1e-9)
    assert np.allclose(u[:, :, 2], 0, atol=1e-12)
