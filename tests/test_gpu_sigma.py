"""The library's Monte Carlo standard errors (batch means over workgroup ray ranges).

Two solves on independent Philox keys must differ by amounts consistent with
√(σ_a² + σ_b²) per output, where both σ come from the library itself: a σ that
is too small fails the |z|>3 count and the χ² upper bound, one that is too
large fails the χ² lower bound.
"""
import numpy as np
import pytest

import mcrt
from mcrt import workloads as W
from statcheck import compare, flat

pytestmark = pytest.mark.gpu


def _flat_sigma(sig):
    return flat((*sig["A"], sig["Irt"], sig["Icb"]))


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_sigma_consistent_with_independent_streams(form):
    cfg, g, seed = W.sphere_mie_cb(65536, Ntheta_table=1801)
    I0 = [1.0, -1.0, 0.0, 0.0]
    print(f"sigma form={form} N={cfg.number_of_rays}", flush=True)
    sa, sb = {}, {}
    a = mcrt.simulate(cfg, g, I0, seed=seed, precision=form, sigma=sa)
    b = mcrt.simulate(cfg, g, I0, seed=seed + 1, precision=form, sigma=sb)
    assert sa["batches"] >= 64 and sa["batches"] == sb["batches"]
    assert sa["Irt"].shape == a[5].shape and sa["Icb"].shape == a[6].shape
    s = np.sqrt(_flat_sigma(sa) ** 2 + _flat_sigma(sb) ** 2)
    assert np.all(np.isfinite(s)) and np.all(s >= 0)
    # the sphere's direct transmission depends on the entry point: σ > 0
    assert sa["A"][0] > 0 and sa["A"][2] > 0
    r = compare(flat(a), flat(b), s, floor=1e-15, dead_rel=1e-6, dof=sa["batches"] - 1,
                label=f"sigma/{form}")
    assert r["ok"], r
    assert r["chi2dof"] > 0.7, r


def test_sigma_needs_cooperative_form():
    cfg, g, seed = W.plane_rayleigh_tau1(256)
    with pytest.raises(mcrt._abi.MCRTError) as e:
        mcrt.simulate(cfg, g, [1.0, 0, 0, 0], seed=seed, precision="fp64_plain", sigma={})
    assert e.value.code == mcrt._abi.MCRT_ERR_UNSUPPORTED
