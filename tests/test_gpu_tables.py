"""Device-resident scatterer tables shared by content (mcrt_table_upload)."""
import numpy as np
import pytest

import mcrt
from mcrt import _abi
from mcrt import workloads as W
from statcheck import flat

pytestmark = pytest.mark.gpu

I0 = (1.0, -1.0, 0.0, 0.0)


def test_pinned_tables_are_found_by_content_and_change_nothing():
    cfg, g, seed = W.sphere_mie_cb(2048, Ntheta_table=361)
    ref = flat(mcrt.simulate(cfg, g, I0, seed=seed, precision="mixed"))
    s1 = mcrt.table_cache_stats()
    with mcrt.ResidentTables(g.scatterer) as t:
        assert t.handle > 0
        s2 = mcrt.table_cache_stats()
        # the plan above left both element sizes in the cache: the upload finds them
        assert s2["hits"] - s1["hits"] == 2 and s2["misses"] == s1["misses"]
        out = flat(mcrt.simulate(cfg, g, I0, seed=seed, precision="mixed"))
        s3 = mcrt.table_cache_stats()
        print(f"[tables] {s1} -> {s2} -> {s3}", flush=True)
        assert s3["hits"] - s2["hits"] == 2 and s3["misses"] == s2["misses"]
        out64 = flat(mcrt.simulate(cfg, g, I0, seed=seed, precision="fp64"))
    assert np.array_equal(out, ref)
    assert np.isfinite(out64).all()
    # a scatterer with other tables is a miss
    cfg2, g2, _ = W.sphere_mie_cb(1024, Ntheta_table=181)
    s4 = mcrt.table_cache_stats()
    with mcrt.ResidentTables(g2.scatterer):
        s5 = mcrt.table_cache_stats()
    assert s5["misses"] - s4["misses"] == 2 and s5["bytes"] > s4["bytes"]


def test_table_handles_are_checked():
    rayleigh = mcrt.RayleighScatterer(1.0, 1.0)
    t = mcrt.ResidentTables(rayleigh)  # no tables, still a handle
    t.release()
    t.release()  # a released handle is not released twice
    ctx = mcrt.get_context()
    with pytest.raises(_abi.MCRTError):
        _abi.check(ctx.lib.mcrt_table_release(ctx.handle, _abi.C.c_uint64(10**12)), ctx.handle)
