"""Device-level known-answer tests of the deterministic building blocks.

Each form's own device functions (mcrt_plan_eval through the C ABI) are
evaluated on fixed inputs and compared value for value with the oracle's FP64
restatement of the reference functions:

  update_I (utils.jl:91-112) as the peel-off evaluates it, update_E
  (utils.jl:170-192) as the CB bins and as the FP64 CB path state evaluate it,
  distance_to_boundary (plane.jl:97-105, sphere.jl:69-73), the phase and
  amplitude lookups (rayleigh.jl:8-23, mie.jl:46-92, spline.jl:26-76 with
  NSPL = Nθ - 1, including P22 and P44), inverse-CDF sampling and
  kepler_solver (utils.jl:243-261).

Bounds: 1e-12 relative to the output's scale in the FP64 forms (fp64_plain
repeats the reference's operations in order; fp64 evaluates the peel-off in
the scattering-plane form, so it agrees to rounding).  The mixed form computes
in FP32: its bound is FP32_REL (1e-5) of the output's scale, where the scale is
max(1, max |reference|) — the FP32 arccos of the table lookup (|error| ≤ 8.8e-8
rad, the degree-6 minimax fit of csrc/fit_acos.py) and the FP32 rounding of μ (≈ 6e-8/sinθ rad,
≲ 2.5e-4 of a 0.1° table step for |μ| ≤ 0.99) dominate; distances are bounded
at FP32_REL of the larger of the medium's size and the distance.  Inputs keep |μ| ≤ 0.99 for Mie lookups,
where the lookup is well conditioned; the polar branch (μ = ±1 exactly) is
exercised with Rayleigh.  At the table nodes themselves the reference's
amplitude lookup is discontinuous (its shifted weights), and either one-sided
value is accepted there, and so it is for the Spline phase lookup, which has
the same weights.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import mcrt
import oracle

pytestmark = pytest.mark.gpu

FORMS = ["fp64_plain", "fp64", "mixed"]
FP64_REL = 1e-12
FP32_REL = 1e-5
rng = np.random.default_rng(77)


def _tol(form):
    return FP32_REL if form == "mixed" else FP64_REL


def _unit(n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def _frames(n):
    t = rng.uniform(0.02, math.pi - 0.02, n)
    p = rng.uniform(0, 2 * math.pi, n)
    return np.stack([np.cos(t), np.sin(t), np.cos(p), np.sin(p)], 1)


def _kdir(f):
    return np.array([f[1] * f[2], f[1] * f[3], f[0]])


def _scatterer(kind):
    if kind == "rayleigh":
        return mcrt.RayleighScatterer(0.9, 1.3)
    if kind == "spline":  # P22 = 0.93·P11, P44 = 0.97·P33 on the Mie tables
        return mcrt.workloads.spline_from(mcrt.workloads.c2_mie(1801))
    return mcrt.workloads.c2_mie(1801)


def _case(kind, geom):
    sca = _scatterer(kind)
    if geom == "plane":
        g = mcrt.PlaneGeometry(0.8 * math.pi, 0.4, [mcrt.Layer(d=2.0 * sca.kl, scatterer=sca, tau=2.0),
                                                      mcrt.Layer(d=1.0 * sca.kl, scatterer=sca, tau=1.0)])
    else:
        g = mcrt.SphereGeometry(R=3.0 * sca.kl, scatterer=sca)
    cfg = mcrt.Config(g, number_of_rays=64, Ntheta=8, Nphi=8, compute_cb=True)
    return sca, g, cfg


def _directions(fr, mie, n):
    """Bin directions k for the frames fr: random, with |μ| ≤ 0.99 for Mie; for
    Rayleigh a few exactly along ±k of the frame (the polar branch)."""
    ks = []
    for i in range(n):
        kin = _kdir(fr[i])
        if not mie and i % 50 == 0:
            ks.append(kin if i % 100 == 0 else -kin)
            continue
        while True:
            k = _unit(1)[0]
            if abs(k @ kin) <= 0.99:
                break
        ks.append(k)
    return np.array(ks)


def _check(form, got, want, label, scale=None, inputs=None):
    got = np.asarray(got, dtype=float)
    want = np.asarray(want, dtype=float)
    sc = np.maximum(1.0, np.abs(want).max(axis=-1)) if scale is None else scale
    err = np.abs(got - want).max(axis=-1) / sc
    worst = float(err.max())
    w = int(err.argmax())
    where = "" if inputs is None else f" at input {np.asarray(inputs)[w]}"
    print(f"kat {form} {label}: max rel err {worst:.3e} (bound {_tol(form):.0e}){where} "
          f"got {got[w]} want {want[w]}")
    assert worst <= _tol(form), (form, label, worst, w)
    return worst


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("kind", ["rayleigh", "mie", "spline"])
def test_kat_update_I_and_E(form, kind):
    sca, g, cfg = _case(kind, "sphere")
    osca = oracle.Sca(sca)
    plan = mcrt.Plan(cfg, g, [1.0, 0.0, 0.0, 0.0], precision=form)
    try:
        n = 600
        fr = _frames(n)
        ks = _directions(fr, kind != "rayleigh", n)
        I = np.concatenate([np.ones((n, 1)), _unit(n) * rng.uniform(0, 1, (n, 1))], 1) * \
            rng.uniform(0.1, 2.0, (n, 1))
        E = rng.normal(size=(n, 4))
        got = plan.eval("update_I", np.concatenate([fr, I, ks], 1))
        want = np.array([oracle.update_I(osca, I[i], ks[i], fr[i])[0] for i in range(n)])
        _check(form, got, want, f"{kind} update_I")
        Ec = E[:, 0::2] + 1j * E[:, 1::2]
        want_e = []
        for i in range(n):
            e, _, _ = oracle.update_E(osca, Ec[i], ks[i], fr[i])
            want_e.append([e[0].real, e[0].imag, e[1].real, e[1].imag])
        want_e = np.array(want_e)
        got = plan.eval("update_E", np.concatenate([fr, E, ks], 1))
        _check(form, got, want_e, f"{kind} update_E (bins)")
        got = plan.eval("update_E_path", np.concatenate([fr, E, ks], 1))
        # the CB path state is FP64 in every form
        _check("fp64", got, want_e, f"{kind} update_E (path state, {form})")
    finally:
        plan.close()


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("geom", ["plane", "sphere"])
def test_kat_distance(form, geom):
    sca, g, cfg = _case("rayleigh", geom)
    S = oracle.Solve(cfg, g)
    plan = mcrt.Plan(cfg, g, [1.0, 0.0, 0.0, 0.0], precision=form)
    try:
        n = 800
        ks = _unit(n)
        if geom == "plane":
            idx = rng.integers(0, 2, n)
            tau = np.where(idx == 0, 2.0, 1.0)
            x = np.stack([rng.uniform(-3, 3, n), rng.uniform(-3, 3, n), -rng.uniform(0, 1, n) * tau], 1)
            size = 2.0
        else:
            idx = np.zeros(n, dtype=int)
            x = _unit(n) * rng.uniform(0, 2.99, (n, 1))
            size = 3.0
        got = plan.eval("distance", np.concatenate([x, ks, idx[:, None]], 1))
        want = np.array([S.distance(x[i], ks[i], int(idx[i])) for i in range(n)])
        fin = np.isfinite(want)
        assert np.array_equal(np.isfinite(got[:, 0]), fin)
        # FP32 keeps a distance to FP32_REL of itself: near-horizontal directions
        # in a slab give distances of thousands of optical depths
        _check(form, got[fin, :1], want[fin, None], f"{geom} distance",
               scale=np.maximum(size, np.abs(want[fin])))
        # the CB path state's FP64 distance
        _check("fp64", got[fin, 1:2], want[fin, None], f"{geom} distance (FP64 path state, {form})",
               scale=size)
    finally:
        plan.close()


def _either_side(got, want, mu, tab, Nt):
    """The reference's weights (msn+1-sn, sn-msn) make a lookup jump at every
    node j: T[j-1] just below, 2T[j] - T[j+1] at and above (mie.jl:76-86,
    spline.jl:35-50).  At a node, rounding of arccos decides the side, so the
    device may take either one-sided value: want takes the one nearer got."""
    sn = np.arccos(mu) / math.pi * (Nt - 1)
    node = np.abs(sn - np.round(sn)) < 1e-9
    j = np.round(sn).astype(int)
    want = want.copy()
    for i in np.nonzero(node)[0]:
        lo = tab[j[i] - 1]
        hi = 2 * tab[j[i]] - tab[j[i] + 1]
        want[i] = lo if np.abs(got[i] - lo).max() < np.abs(got[i] - hi).max() else hi
    return want, int(node.sum())


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("kind", ["rayleigh", "mie", "spline"])
def test_kat_lookups_sampling_kepler(form, kind):
    sca, g, cfg = _case(kind, "sphere")
    osca = oracle.Sca(sca)
    plan = mcrt.Plan(cfg, g, [1.0, 0.0, 0.0, 0.0], precision=form)
    try:
        n = 2000
        mu = rng.uniform(-0.99, 0.99, n)
        if kind != "rayleigh":  # table nodes and midpoints too
            th = (np.arange(0, 1800, 9) + rng.choice([0.0, 0.5], 200)) * math.pi / 1800
            mu = np.concatenate([mu, np.cos(th[np.abs(np.cos(th)) <= 0.99])])
        else:
            mu = np.concatenate([mu, [-1.0, 1.0, 0.0]])
        P6 = np.array([oracle.phase_eval(osca, m) for m in mu])  # P11 P12 P22 P33 P34 P44
        P = P6[:, [0, 1, 3, 4]]
        D = P6[:, [2, 5, 2, 5]]
        for what, want, label in (("phase", P, "phase (peel-off)"),
                                  ("phase_transport", P, "phase (transport)"),
                                  ("phase_diag", D, "P22, P44 (peel-off, transport)")):
            got = plan.eval(what, mu[:, None])
            if kind == "spline":
                cols = {"phase": [0, 1, 3, 4], "phase_transport": [0, 1, 3, 4],
                        "phase_diag": [2, 5, 2, 5]}[what]
                tab = np.stack([sca.F11, sca.F12, sca.F22, sca.F33, sca.F34, sca.F44], 1)[:, cols]
                # each lookup takes its own side of a node: phase_diag holds two
                groups = [[0, 1], [2, 3]] if what == "phase_diag" else [[0, 1, 2, 3]]
                want = want.copy()
                for gcols in groups:
                    w, nn = _either_side(got[:, gcols], want[:, gcols], mu, tab[:, gcols],
                                         sca.Ntheta)
                    want[:, gcols] = w
                    assert nn > 50  # the node cases were exercised
            _check(form, got, want, f"{kind} {label}", inputs=mu)
        S = np.array([[v.real, v.imag, w.real, w.imag] for v, w in (oracle.amplitude(osca, m) for m in mu)])
        got = plan.eval("amplitude", mu[:, None])
        if kind != "rayleigh":
            tab = np.stack([sca.S11.real, sca.S11.imag, sca.S22.real, sca.S22.imag], 1)
            S, nn = _either_side(got, S, mu, tab, sca.Ntheta)
            assert nn > 50  # the node cases were exercised
        _check(form, got, S, f"{kind} amplitude", inputs=mu)
        u = np.concatenate([rng.uniform(1e-6, 1 - 1e-6, 1000), [0.5, 0.25, 0.75]])
        want = np.array([oracle.phase_sample(osca, x) for x in u])
        _check(form, plan.eval("sample", u[:, None])[:, :1], want[:, None], f"{kind} sample")
        M = rng.uniform(0, 6 * math.pi, 300)
        e = rng.uniform(0, 0.99, 300)
        want = np.array([oracle.kepler(a, b) for a, b in zip(M, e)])
        got = plan.eval("kepler", np.stack([M, e], 1))[:, :1]
        if form == "mixed":  # FP32 Newton-type stop at 2e-6 (mcrt_device.hpp), FP32 M
            assert np.max(np.abs(got[:, 0] - want) / np.maximum(1, np.abs(want))) < 1e-4
        else:
            _check(form, got, want[:, None], "kepler", scale=np.maximum(1, np.abs(want)))
    finally:
        plan.close()


@pytest.mark.parametrize("form", ["fp64", "mixed", "fp64_plain"])
def test_kat_bin_math_and_cb_phase(form):
    """The bin loops' own e^{-t}, √x, 1/y and table arccos (BIN_MATH) and the CB
    interference phase's cos/sin of 2πr (CB_PHASE) against 30-digit values:
    within a few ulp in the FP64 form (branch-free polynomials and Newton steps
    instead of the library functions), within FP32 rounding in the mixed form
    (hardware exp2/sqrt/rcp/sin/cos, degree-6 arccos).  The plain form has no
    such loops and refuses both."""
    import mpmath as mp

    sca, g, cfg = _case("mie", "sphere")
    plan = mcrt.Plan(cfg, g, [1.0, 0.0, 0.0, 0.0], precision=form)
    try:
        n = 1500
        t = np.concatenate([rng.uniform(0.0, 40.0, n - 4), [0.0, 1e-300, 1e-8, 30.0]])
        x = np.concatenate([10.0 ** rng.uniform(-30, 4, n - 2), [0.0, 1.0]])
        y = np.concatenate([10.0 ** rng.uniform(-24, 0.6, n - 1), [1.0]])
        lim = 0.99 if form == "mixed" else 1.0  # FP32 μ near ±1: arccos ill-conditioned
        mu = np.concatenate([rng.uniform(-lim, lim, n - 3), [lim, -lim, 0.0]])
        r = np.concatenate([rng.uniform(-3000.0, 3000.0, 700), rng.uniform(-1.0, 1.0, 700),
                            [0.0, 0.125, 0.25, 0.5, 0.75, -0.25, -0.5, 1.0]])
        if form == "fp64_plain":
            for what, rec in (("bin_math", np.stack([t, x, y, mu], 1)), ("cb_phase", r[:, None])):
                with pytest.raises(mcrt._abi.MCRTError) as e:
                    plan.eval(what, rec)
                assert e.value.code == mcrt._abi.MCRT_ERR_UNSUPPORTED
            return
        mp.mp.dps = 30
        got = plan.eval("bin_math", np.stack([t, x, y, mu], 1))
        want = np.array([[float(mp.exp(-mp.mpf(a))), float(mp.sqrt(mp.mpf(b))),
                          float(1 / mp.mpf(c)), float(mp.acos(mp.mpf(d)))]
                         for a, b, c, d in zip(t, x, y, mu)])
        rel = np.abs(got[:, :3] - want[:, :3]) / np.where(want[:, :3] == 0, 1.0, np.abs(want[:, :3]))
        acos_err = np.abs(got[:, 3] - want[:, 3])
        bounds = (1e-15, 1e-15, 1e-15, 4e-15) if form == "fp64" else (1e-5, 1e-6, 1e-6, 2e-6)
        for c, (label, err) in enumerate((("exp(-t)", rel[:, 0]), ("sqrt", rel[:, 1]),
                                          ("1/y", rel[:, 2]), ("acos", acos_err))):
            w = int(np.argmax(err))
            print(f"kat {form} bin {label}: max err {err[w]:.3e} (bound {bounds[c]:.0e}) "
                  f"at {(t[w], x[w], y[w], mu[w])[c]!r}")
            assert err[w] <= bounds[c], (form, label, err[w])
        assert np.all(got[x == 0.0, 1] == 0.0)
        got = plan.eval("cb_phase", r[:, None])
        want = np.array([[float(mp.cos(2 * mp.pi * mp.mpf(v))), float(mp.sin(2 * mp.pi * mp.mpf(v)))]
                         for v in r])
        err = np.abs(got[:, :2] - want).max(axis=1)
        w = int(np.argmax(err))
        bound = 2e-15 if form == "fp64" else 1e-5
        print(f"kat {form} CB phase cos/sin: max abs err {err[w]:.3e} (bound {bound:.0e}) at r = {r[w]!r}")
        assert err[w] <= bound, (form, err[w])
    finally:
        plan.close()
