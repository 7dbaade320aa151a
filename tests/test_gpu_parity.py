"""GPU parity: every precision form against the oracle on the same Philox stream.

The oracle (oracle/liboracle.so) restates the reference CPU path; the GPU
forms consume the same counter-based draws, so their per-ray histories agree
with it up to rounding.  Acceptance is statistical (binomial count of |z|>3
plus χ²/dof, with σ from the oracle's batch means), plus exact energy
telescoping.  Sizes are chosen so the oracle finishes in seconds.
"""
import numpy as np
import pytest

import mcrt
import oracle
from mcrt import workloads as W
from statcheck import compare, flat, oracle_flat

pytestmark = pytest.mark.gpu

FORMS = ["fp64_plain", "fp64", "mixed"]


def _cases():
    return {
        "plane_rayleigh": W.plane_rayleigh_tau1(2048),
        "plane_rayleigh_cb": W.plane_rayleigh_tau1(1024, compute_cb=True),
        "sphere_mie_cb": W.sphere_mie_cb(512, Ntheta_table=1801),
    }


_ORACLE = {}


def _oracle(name, cfg, g, seed, I0):
    key = (name, tuple(I0))
    if key not in _ORACLE:
        _ORACLE[key] = oracle.Solve(cfg, g).simulate(I0, seed, nbatch=32)
    return _ORACLE[key]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("case", ["plane_rayleigh", "plane_rayleigh_cb", "sphere_mie_cb"])
def test_form_vs_oracle(case, form):
    cfg, g, seed = _cases()[case]
    I0 = [1.0, -1.0, 0.0, 0.0]
    print(f"case={case} form={form} N={cfg.number_of_rays}", flush=True)
    ref = _oracle(case, cfg, g, seed, I0)
    st = {}
    try:
        out = mcrt.simulate(cfg, g, I0, seed=seed, precision=form, stats=st)
    except mcrt._abi.MCRTError as e:  # form not built yet
        if e.code == mcrt._abi.MCRT_ERR_UNSUPPORTED:
            pytest.skip(str(e))
        raise
    a = flat(out)
    b = oracle_flat(ref)
    sig = oracle.sigma_of_mean(ref) * np.sqrt(2.0)
    # energy budget telescopes to 1 per ray
    assert abs(sum(out[:5]) - 1.0) < (1e-5 if form == "mixed" else 1e-9)
    assert st["rays"] == cfg.number_of_rays
    assert st["nonfinite"] == 0
    # deterministic values (σ = 0, e.g. the direct transmission e^{-τ'}) agree to
    # FP64 rounding in the FP64 forms and to FP32 rounding in the mixed form
    dead_rel = 1e-6 if form == "mixed" else 1e-12
    r = compare(a, b, sig, floor=1e-15, dead_rel=dead_rel, label=f"{case}/{form}")
    assert r["ok"], r
    # same stream: the per-ray histories agree, so every bin agrees far inside σ
    assert r["max_abs_z"] < (1.0 if form == "mixed" else 0.5), r
    # same stream: event counts agree closely
    ev_ref = ref["counters"][0]
    assert abs(st["events"] - ev_ref) <= 0.02 * ev_ref + 5


_BIG = []


@pytest.mark.parametrize("form", FORMS)
def test_headline_workload_at_16384_rays(form):
    """The bench's own workload (sphere τ = 5 of Mie spheres, Mie table of 1801
    nodes, CB on) at 16384 rays of one stream in every form against the oracle:
    32× the rays of the case above, so each bin's σ is ~6× smaller and the same
    same-stream criteria are that much tighter in absolute terms."""
    cfg, g, seed = W.sphere_mie_cb(16384, Ntheta_table=1801)
    I0 = [1.0, -1.0, 0.0, 0.0]
    print(f"case=sphere_mie_cb form={form} N={cfg.number_of_rays}", flush=True)
    if not _BIG:
        _BIG.append(oracle.Solve(cfg, g).simulate(I0, seed, nbatch=32))
    ref = _BIG[0]
    st = {}
    out = mcrt.simulate(cfg, g, I0, seed=seed, precision=form, stats=st)
    assert abs(sum(out[:5]) - 1.0) < (1e-5 if form == "mixed" else 1e-9)
    assert st["rays"] == cfg.number_of_rays and st["nonfinite"] == 0
    r = compare(flat(out), oracle_flat(ref), oracle.sigma_of_mean(ref) * np.sqrt(2.0),
                floor=1e-15, dead_rel=1e-6 if form == "mixed" else 1e-12,
                label=f"sphere_mie_cb 16384/{form}")
    assert r["ok"], r
    assert r["max_abs_z"] < (1.0 if form == "mixed" else 0.5), r
    ev_ref = ref["counters"][0]
    assert abs(st["events"] - ev_ref) <= 0.02 * ev_ref + 5


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_independent_streams_agree_within_3sigma(form):
    """A genuinely statistical check: the GPU form on one Philox key against the
    oracle on another key; the difference of two independent estimates must be
    consistent with √2 σ per bin (binomial |z|>3 count and χ²/dof)."""
    cfg, g, seed = W.sphere_mie_cb(2048, Ntheta_table=1801)
    I0 = [1.0, -1.0, 0.0, 0.0]
    ref = oracle.Solve(cfg, g).simulate(I0, seed + 1000, nbatch=32)
    out = mcrt.simulate(cfg, g, I0, seed=seed, precision=form)
    sig = oracle.sigma_of_mean(ref) * np.sqrt(2.0)
    r = compare(flat(out), oracle_flat(ref), sig, floor=1e-15, dead_rel=1e-6, dof=31,
                label=f"indep/{form}")
    assert r["ok"], r
    assert r["chi2dof"] > 0.5  # the streams really are independent


@pytest.mark.parametrize("form", FORMS)
def test_ray_range_matches_oracle(form):
    """Rays beyond 2^32 of the stream: the GPU plan over [r0, r0 + N) against the
    oracle over the same counter range (same-stream parity, as above)."""
    cfg, g, seed = W.plane_rayleigh_tau1(1024, compute_cb=True)
    I0 = [1.0, -1.0, 0.0, 0.0]
    r0, inst = 3_000_000_017, 5
    print(f"range form={form} r0={r0} instance={inst}", flush=True)
    ref = oracle.Solve(cfg, g).simulate(I0, seed, ray_begin=r0, instance=inst, nbatch=32)
    st = {}
    out = mcrt.simulate(cfg, g, I0, seed=seed, precision=form, stats=st, ray_begin=r0,
                        instance=inst)
    sig = oracle.sigma_of_mean(ref) * np.sqrt(2.0)
    dead_rel = 1e-6 if form == "mixed" else 1e-12
    r = compare(flat(out), oracle_flat(ref), sig, floor=1e-15, dead_rel=dead_rel,
                label=f"range/{form}")
    assert r["ok"], r
    assert r["max_abs_z"] < (1.0 if form == "mixed" else 0.5), r
    assert abs(st["events"] - ref["counters"][0]) <= 0.02 * ref["counters"][0] + 5
    # a different stream position really gives different rays
    base = mcrt.simulate(cfg, g, I0, seed=seed, precision=form)
    assert not np.allclose(flat(base)[5:], flat(out)[5:])


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_shards_reassemble_the_solve(form):
    """Two plans over disjoint halves of one stream average to the whole solve
    (the basis of sharding rays across devices)."""
    cfg, g, seed = W.sphere_mie_cb(8192, Ntheta_table=1801)
    I0 = [1.0, -1.0, 0.0, 0.0]
    whole = flat(mcrt.simulate(cfg, g, I0, seed=seed, precision=form))
    half = cfg.replace(number_of_rays=4096)
    a = flat(mcrt.simulate(half, g, I0, seed=seed, precision=form, ray_begin=0))
    b = flat(mcrt.simulate(half, g, I0, seed=seed, precision=form, ray_begin=4096))
    merged = 0.5 * (a + b)
    scale = np.max(np.abs(whole))
    tol = (1e-6 if form == "mixed" else 1e-12) * scale
    err = float(np.max(np.abs(merged - whole)))
    print(f"[shards/{form}] max|Δ|={err:.3e} tol={tol:.3e}", flush=True)
    assert err <= tol


def test_mixed_cb_stays_finite_at_scale():
    """1e7 rays of the plane CB case in the mixed form: FP32 rounding of fully
    polarised rays makes the backscattered intensity of some events ≤ 0; such
    events add nothing (the reference would raise DomainError or give NaN)."""
    cfg, g, seed = W.plane_rayleigh_cb(10**7)
    st = {}
    out = mcrt.simulate(cfg, g, [1.0, -1.0, 0.0, 0.0], seed=seed, precision="mixed", stats=st)
    a = flat(out)
    print(f"[finite/mixed] nonfinite outputs {int((~np.isfinite(a)).sum())} rays "
          f"{st['rays']} nonfinite rays {st['nonfinite']}", flush=True)
    assert np.all(np.isfinite(a))
    assert st["nonfinite"] == 0
