"""Batched (one persistent launch) and served solves against the oracle.

Sixteen instances span the corners of the batched population (slab and sphere,
Rayleigh and Mie, optical thickness or radius 0.1 to 20, size parameter 0.5
to 20, Im m 1e-5 to 1e-1, Re m 1.3 to 1.7), with CB on the population's
grid (θb 0–5° every 0.05° plus 6–180° every 1°, × 8 azimuths) and 180 × 48
peel-off bins, 512 to 2048 rays each, half of them in the mixed form and half
in FP64; four more mix slabs whose layers hold different scatterers and
SplineScatterer solves into a batch and a server.  Each solve has its own Philox stream (seed, instance).  Every batched
result and every served result (all sixteen submitted at once, so they share
launches) is compared with the oracle on that same stream, per bin, with σ
from the oracle's 32 batch means: the |z| > 3 count under a binomial bound
and χ²/dof (statcheck.compare with √2σ), the energy budget, and, because the
per-ray histories are the same, every bin within 1σ (mixed) or 0.5σ (FP64) as
in the single-solve parity tests.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import mcrt
import oracle
from mcrt.batch import BATCH_THETAB_DEG, Batch, Request
from mcrt.config import midpoint_polar_grid
from mcrt.serve import Server
from statcheck import compare, flat, oracle_flat

pytestmark = pytest.mark.gpu

I0 = (1.0, -1.0, 0.0, 0.0)
LAM, NU = 0.55, 0.1

# (geometry, scatterer, τ, x, m, rays)
CORNERS = [
    ("plane", "rayleigh", 0.1, None, None, 2048),
    ("plane", "rayleigh", 20.0, None, None, 512),
    ("sphere", "rayleigh", 0.1, None, None, 2048),
    ("sphere", "rayleigh", 20.0, None, None, 512),
    ("plane", "mie", 0.1, 0.5, 1.3 + 1e-5j, 2048),
    ("plane", "mie", 20.0, 20.0, 1.7 + 1e-5j, 512),
    ("plane", "mie", 20.0, 0.5, 1.5 + 1e-1j, 512),
    ("plane", "mie", 0.1, 20.0, 1.3 + 1e-1j, 2048),
    ("sphere", "mie", 0.1, 0.5, 1.7 + 1e-1j, 2048),
    ("sphere", "mie", 20.0, 20.0, 1.3 + 1e-5j, 512),
    ("sphere", "mie", 20.0, 0.5, 1.5 + 1e-5j, 512),
    ("sphere", "mie", 0.1, 20.0, 1.7 + 1e-5j, 2048),
    ("plane", "mie", 2.0, 5.7, 1.3 + 1e-3j, 1024),
    ("sphere", "mie", 5.0, 10.0, 1.7 + 1e-2j, 1024),
    ("plane", "mie", 1.0, 15.0, 1.5 + 1e-1j, 1024),
    ("sphere", "rayleigh", 2.0, None, None, 1024),
]


def _instances():
    out = []
    for i, (geo, kind, tau, x, m, n) in enumerate(CORNERS):
        if kind == "rayleigh":
            sca = mcrt.RayleighScatterer(1.0, 1.0)
        else:
            sca = mcrt.MieScatterer.create(x * LAM / (2 * math.pi), m, LAM, NU, Ntheta=1801,
                                           random_points=1000)
        if geo == "plane":
            g = mcrt.PlaneGeometry(math.pi, 0.0, [mcrt.Layer(d=tau * sca.kl, scatterer=sca, tau=tau,
                                                             omega=sca.omega)])
        else:
            g = mcrt.SphereGeometry(R=tau * sca.kl, scatterer=sca)
        ct, st, w = midpoint_polar_grid(g, 180, 48)
        cfg = mcrt.Config(g, number_of_rays=n, maximum_scattering_times=10000, compute_cb=True,
                          Ntheta=180, Nphi=48, costheta=ct, sintheta=st, inorm=w,
                          thetab=np.deg2rad(BATCH_THETAB_DEG), Nphib=8, minimum_intensity=1e-6,
                          tau0=50.0)
        form = "mixed" if i % 2 == 0 else "fp64"
        out.append(Request(cfg, g, I0, seed=9000 + i, precision=form, instance=100 + i,
                           meta=dict(index=i, label=f"{geo}/{kind}/tau={tau}/x={x}/m={m}")))
    return out


_REF: dict = {}


def _oracle(r: Request):
    key = r.meta["index"]
    if key not in _REF:
        _REF[key] = oracle.Solve(r.cfg, r.g).simulate(list(r.I0), r.seed, nbatch=32,
                                                       instance=r.instance)
    return _REF[key]


def _check_vs_oracle(r: Request, out, st, where):
    ref = _oracle(r)
    form = r.precision
    label = f"{where}/{r.meta['index']}/{form}/{r.meta['label']}"
    print(f"{label}: rays={st['rays']} events={st['events']} (oracle {ref['counters'][0]})", flush=True)
    assert st["rays"] == r.cfg.number_of_rays and st["nonfinite"] == 0, (label, st)
    a = flat(out)
    assert abs(a[:5].sum() - 1.0) < (1e-5 if form == "mixed" else 1e-9), (label, a[:5])
    sig = oracle.sigma_of_mean(ref) * math.sqrt(2.0)
    res = compare(a, oracle_flat(ref), sig, floor=1e-15, dead_rel=1e-6 if form == "mixed" else 1e-12,
                  label=label)
    assert res["ok"], (label, res)
    assert res["max_abs_z"] < (1.0 if form == "mixed" else 0.5), (label, res)
    ev = ref["counters"][0]
    assert abs(st["events"] - ev) <= 0.02 * ev + 5, (label, st["events"], ev)


@pytest.mark.timeout(400)
def test_batched_population_corners_match_the_oracle():
    reqs = _instances()
    with Batch(reqs) as b:
        b.launch()
        b.wait()
        print(f"batch {b.info()} {b.kernel_ms():.1f} ms", flush=True)
        for i, r in enumerate(reqs):
            st = {}
            _check_vs_oracle(r, b.read(i, stats=st), st, "batch")


def _layered_and_spline():
    """Solves that need per-layer tables or the Spline lookup, beside two corners."""
    mie = mcrt.MieScatterer.create(0.3, complex(1.45, 0.01), LAM, NU, Ntheta=721, random_points=1000)
    sp = mcrt.workloads.spline_from(mie)
    ray = mcrt.RayleighScatterer(1.0, 1.0)
    geos = [
        ("rayleigh-over-mie slab", mcrt.PlaneGeometry(math.pi, 0.0, [
            mcrt.Layer(d=0.5, scatterer=ray, tau=0.5, omega=1.0),
            mcrt.Layer(d=2.0 * mie.kl, scatterer=mie, tau=2.0, omega=mie.omega)])),
        ("mie-over-semi-infinite-rayleigh slab", mcrt.PlaneGeometry(math.radians(140.0), 0.3, [
            mcrt.Layer(d=1.0 * mie.kl, scatterer=mie, tau=1.0, omega=mie.omega),
            mcrt.Layer(d=-1.0, scatterer=mcrt.RayleighScatterer(0.95, 1.0), tau=1.0, omega=0.95)])),
        ("spline sphere", mcrt.SphereGeometry(R=3.0 * sp.kl, scatterer=sp)),
        ("rayleigh-over-spline slab", mcrt.PlaneGeometry(math.pi, 0.0, [
            mcrt.Layer(d=0.3, scatterer=ray, tau=0.3, omega=1.0),
            mcrt.Layer(d=1.5 * sp.kl, scatterer=sp, tau=1.5, omega=sp.omega)])),
    ]
    out = []
    for i, (label, g) in enumerate(geos):
        ct, st, w = midpoint_polar_grid(g, 90, 24)
        cfg = mcrt.Config(g, number_of_rays=1024, maximum_scattering_times=10000, compute_cb=True,
                          Ntheta=90, Nphi=24, costheta=ct, sintheta=st, inorm=w,
                          thetab=np.deg2rad(BATCH_THETAB_DEG[::4]), Nphib=8,
                          minimum_intensity=1e-6, tau0=50.0)
        out.append(Request(cfg, g, I0, seed=9500 + i, precision="mixed" if i % 2 == 0 else "fp64",
                           instance=200 + i, meta=dict(index=100 + i, label=label)))
    corners = _instances()
    return out + [corners[5], corners[10]]


@pytest.mark.timeout(400)
def test_batched_and_served_layered_and_spline_solves_match_the_oracle():
    reqs = _layered_and_spline()
    with Batch(reqs) as b:
        b.launch()
        b.wait()
        for i, r in enumerate(reqs):
            st = {}
            _check_vs_oracle(r, b.read(i, stats=st), st, "batch")
    with Server() as srv:
        tickets = [srv.submit(r) for r in reqs]
        for t, r in zip(tickets, reqs):
            st = {}
            _check_vs_oracle(r, srv.result(t, stats=st), st, "served")


@pytest.mark.timeout(400)
def test_served_population_corners_match_the_oracle():
    reqs = _instances()
    with Server() as srv:
        tickets = [srv.submit(r) for r in reqs]  # all in flight together
        outs = []
        for t, r in zip(tickets, reqs):
            st = {}
            outs.append((r, srv.result(t, stats=st), st))
        s = srv.stats()
    print(f"server {s}", flush=True)
    assert s["completed"] == len(reqs) and s["failed"] == 0
    for r, out, st in outs:
        _check_vs_oracle(r, out, st, "served")
