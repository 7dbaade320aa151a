"""The request loop: continuous batching, alone-vs-mix equality, cancellation,
admission by device memory, the Mie table cache."""
import numpy as np
import pytest

import mcrt
from mcrt import workloads as W
from mcrt.batch import Batch, Request, random_instances
from mcrt.serve import Server, poisson_load
from statcheck import flat

pytestmark = pytest.mark.gpu

I0 = (1.0, -1.0, 0.0, 0.0)


def _reqs():
    a = W.plane_rayleigh_tau1(4096, compute_cb=True)
    b = W.sphere_mie_cb(4096, Ntheta_table=1801)
    c = W.sphere_mie_cb(2048, Ntheta_table=1801)
    return [Request(a[0], a[1], I0, seed=11, precision="mixed"),
            Request(b[0], b[1], I0, seed=12, precision="mixed"),
            Request(c[0], c[1], I0, seed=13, precision="fp64"),
            Request(b[0], b[1], I0, seed=14, precision="mixed")]


def test_result_is_the_same_alone_and_in_the_mix():
    reqs = _reqs()
    with Server() as srv:
        t = srv.submit(reqs[1])
        assert srv.poll(t) in ("queued", "running", "done")
        assert srv.wait(t) == "done" and srv.poll(t) == "done"
        alone = flat(srv.result(t))
        ts = [srv.submit(r) for r in reqs]  # reqs[1] again, now among others
        outs = [flat(srv.result(x)) for x in ts]
        st = srv.stats()
    print(f"[serve] stats {st}", flush=True)
    err = float(np.max(np.abs(outs[1] - alone)))
    print(f"[serve/alone-vs-mix] max|Δ|={err:.3e}", flush=True)
    assert err <= 1e-12 * float(np.max(np.abs(alone)))
    # and, cut into the same chunks, the same as the batched launch of that
    # request (batches and the server have different default chunk sizes)
    r = reqs[1]
    fixed = Request(r.cfg, r.g, r.I0, seed=r.seed, precision=r.precision, chunk_rays=512)
    with Server() as srv:
        served = flat(srv.result(srv.submit(fixed)))
    with Batch([fixed]) as b:
        b.launch()
        b.wait()
        ref = flat(b.read(0))
    assert float(np.max(np.abs(ref - served))) <= 1e-12 * float(np.max(np.abs(ref)))
    for r, o in zip(reqs, outs):
        assert np.all(np.isfinite(o)) and abs(o[:5].sum() - 1.0) < 1e-5
    assert st["completed"] == 5 and st["failed"] == 0
    assert st["table_hits"] >= 1  # the Mie tables of the sphere stayed resident


def test_cancel_and_continue():
    cfg, g, seed = W.sphere_mie_cb(2_000_000, Ntheta_table=1801)
    small = W.plane_rayleigh_tau1(2048)
    with Server() as srv:
        big = srv.submit(Request(cfg, g, I0, seed=1, precision="mixed"))
        srv.cancel(big)
        state = srv.wait(big)
        assert state in ("cancelled", "done")
        if state == "cancelled":
            with pytest.raises(mcrt._abi.MCRTError) as e:
                srv.result(big)
            assert e.value.code == mcrt._abi.MCRT_ERR_CANCELLED
        t = srv.submit(Request(small[0], small[1], I0, seed=2, precision="fp64"))
        out = flat(srv.result(t))
        assert abs(out[:5].sum() - 1.0) < 1e-9
        st = srv.stats()
    print(f"[serve/cancel] {state} {st}", flush=True)
    assert st["cancelled"] + st["completed"] == 2


def test_admission_by_device_memory():
    reqs = _reqs()[:2]
    from mcrt import _abi  # noqa: F401
    # a budget that holds one solve at a time
    with Server(device_budget=40 << 20) as srv:
        ts = [srv.submit(r) for r in reqs * 2]
        outs = [flat(srv.result(t)) for t in ts]
        st = srv.stats()
    print(f"[serve/admission] {st}", flush=True)
    assert st["completed"] == 4
    assert float(np.max(np.abs(outs[0] - outs[2]))) <= 1e-12 * float(np.max(np.abs(outs[0])))


def test_poisson_load_small():
    pool = random_instances(16, seed=3, ray_scale=1e-4)
    with Server() as srv:
        rep = poisson_load(srv, pool, rate=50.0, duration=2.0, seed=9)
    rep.pop("kept")
    print(f"[serve/load] {rep}", flush=True)
    assert rep["completed"] == rep["requests"] and rep["errors"] == 0
    assert rep["p99_ms"] >= rep["p50_ms"] > 0


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_result_is_the_same_when_windows_are_cut(precision):
    """A solve whose windows are stopped again and again by arriving work
    resumes its chunks exactly where they stopped: the result equals the
    batched launch of the same chunks up to the order of FP64 additions."""
    cfg, g, _ = W.sphere_mie_cb(32768 if precision == "mixed" else 8192, Ntheta_table=1801)
    big = Request(cfg, g, I0, seed=21, precision=precision, chunk_rays=512)
    small = W.plane_rayleigh_tau1(1024, compute_cb=True)
    import time
    with Server() as srv:
        srv.set_slice(0.0)  # stop the running window whenever work waits
        tb = srv.submit(big)
        ts = []
        for i in range(24):
            time.sleep(0.004)
            ts.append(srv.submit(Request(small[0], small[1], I0, seed=100 + i, precision=precision)))
        served = flat(srv.result(tb))
        for t in ts:
            srv.result(t)
        st = srv.stats()
    print(f"[serve/cut {precision}] {st}", flush=True)
    with Batch([big]) as b:
        b.launch()
        b.wait()
        ref = flat(b.read(0))
    err = float(np.max(np.abs(ref - served)))
    print(f"[serve/cut {precision}] max|Δ| = {err:.3e}", flush=True)
    assert st["preemptions"] > 0 and st["resumed"] > 0
    assert st["completed"] == 25 and st["failed"] == 0
    assert err <= 1e-12 * float(np.max(np.abs(ref)))
