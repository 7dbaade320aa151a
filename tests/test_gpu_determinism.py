"""Bitwise reproducibility of the cooperative forms.

Every sum of a solve is taken in a fixed order: escape sums over the bins of a
wave (DPP) and over waves, bin sums in the registers of the one thread that owns
the bin, workgroup slabs summed in workgroup order.  So a plan launched twice,
and a second plan over the same stream, give identical outputs.  The cases have
more than 64 events per iteration (escape sums in several blocks of 64) and
iterations with few events (the ray tail).
"""
import numpy as np
import pytest

import mcrt
from mcrt import workloads as W
from statcheck import flat

pytestmark = pytest.mark.gpu

I0 = [1.0, -1.0, 0.0, 0.0]


def _case(name):
    if name == "plane_rayleigh_cb":
        return W.plane_rayleigh_tau1(4096, compute_cb=True)
    return W.sphere_mie_cb(2048, Ntheta_table=361)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["plane_rayleigh_cb", "sphere_mie_cb"])
def test_relaunch_and_second_plan_are_bitwise_identical(case, form):
    cfg, g, seed = _case(case)
    print(f"determinism case={case} form={form} N={cfg.number_of_rays}", flush=True)
    p = mcrt.Plan(cfg, g, I0, seed=seed, precision=form)
    try:
        p.launch()
        st = {}
        a = flat(p.read(stats=st))
        p.launch()
        b = flat(p.read())
    finally:
        p.close()
    q = mcrt.Plan(cfg, g, I0, seed=seed, precision=form)
    try:
        q.launch()
        c = flat(q.read())
    finally:
        q.close()
    assert st["rays"] == cfg.number_of_rays and st["nonfinite"] == 0
    assert np.all(np.isfinite(a))
    assert np.array_equal(a, b)
    assert np.array_equal(a, c)
