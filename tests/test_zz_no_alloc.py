"""mcrt_simulate_static (no device allocation) against the entry points it varies.

Two calls in a row through the C ABI, the second with a smaller solve (fewer
rays and bins), each compared with
  * a one-request batch of the same solve (mcrt_batch_*: the same chunks and
    kernels, so equal up to the order of FP64 additions, 1e-12 relative), and
  * mcrt_simulate on the same Philox stream (the cooperative plan: per-ray
    histories agree, so every output within 1σ of its batch-means σ in the
    mixed form and 0.5σ in FP64, with the |z| > 3 count and χ²/dof bounds of
    the parity tests).
A solve whose device state exceeds the fixed scratch array is refused with
MCRT_ERR_NOMEM before anything runs.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import mcrt
from mcrt import _abi
from mcrt import workloads as W
from mcrt.api import _shape_outputs, get_context, marshal
from mcrt.batch import Batch, Request
from statcheck import compare, flat

pytestmark = pytest.mark.gpu

I0 = [1.0, -1.0, 0.0, 0.0]


def _static(cfg, g, seed, form):
    ctx = get_context(0)
    cs, gs, i0, keep = marshal(cfg, g, I0)
    A = np.full(5, np.nan)
    Irt = np.full(4 * cfg.Ntheta * cfg.Nphi, np.nan)
    Icb = np.full(max(1, 4 * cfg.Nthetab * cfg.Nphib), np.nan)
    st = _abi.mcrt_stats()
    rc = ctx.lib.mcrt_simulate_static(ctx.handle, C.byref(cs), C.byref(gs), _abi.dptr(i0),
                                      C.c_uint64(seed), _abi.PRECISIONS[form], _abi.dptr(A),
                                      _abi.dptr(Irt), _abi.dptr(Icb), C.byref(st))
    return rc, _shape_outputs(cfg, A, Irt, Icb[: 4 * cfg.Nthetab * cfg.Nphib]), st.as_dict()


def _check(cfg, g, seed, form, label):
    print(f"static {label} form={form} N={cfg.number_of_rays}", flush=True)
    rc, out, st = _static(cfg, g, seed, form)
    assert rc == 0, (rc, _abi.load_library().mcrt_last_error(None))
    a = flat(out)
    assert np.all(np.isfinite(a)) and st["rays"] == cfg.number_of_rays and st["nonfinite"] == 0
    assert abs(a[:5].sum() - 1.0) < (1e-5 if form == "mixed" else 1e-9)
    # a one-request batch of the same solve
    with Batch([Request(cfg, g, tuple(I0), seed=seed, precision=form)]) as b:
        b.launch()
        b.wait()
        bst = {}
        ref = flat(b.read(0, stats=bst))
    err = float(np.max(np.abs(a - ref)))
    print(f"[static/{label}/{form}] vs batch max|Δ|={err:.3e}", flush=True)
    assert err <= 1e-12 * float(np.max(np.abs(ref)))
    assert st["events"] == bst["events"] and st["chunks"] == bst["chunks"]
    # mcrt_simulate on the same stream, with its batch-means σ
    sig = {}
    plan = flat(mcrt.simulate(cfg, g, I0, seed=seed, precision=form, sigma=sig))
    s = flat((*sig["A"], sig["Irt"], sig["Icb"]))
    r = compare(a, plan, s, floor=1e-15, dead_rel=1e-6 if form == "mixed" else 1e-12,
                label=f"static-vs-simulate/{label}/{form}")
    assert r["ok"], r
    assert r["max_abs_z"] < (1.0 if form == "mixed" else 0.5), r
    return a


@pytest.mark.parametrize("form", ["mixed", "fp64"])
def test_static_matches_simulate_and_batch_twice(form):
    big = W.sphere_mie_cb(65536 if form == "mixed" else 16384, Ntheta_table=1801)
    small = W.plane_rayleigh_tau1(4096, compute_cb=True, Ntheta=40, Nphi=16)
    first = _check(big[0], big[1], big[2], form, "sphere_mie_cb")
    second = _check(small[0], small[1], small[2], form, "plane_rayleigh_cb (smaller)")
    assert first.shape != second.shape


def test_static_refuses_solves_beyond_its_scratch():
    # 1e6 rays cut into 123 chunks: 123 batch slabs of 1.04e6 outputs (about 1 GB)
    cfg, g, seed = W.plane_rayleigh_tau1(10**6, compute_cb=True, Ntheta=720, Nphi=360)
    rc, _, _ = _static(cfg, g, seed, "mixed")
    msg = _abi.load_library().mcrt_last_error(get_context(0).handle).decode()
    print(f"[static/too-big] rc={rc} {msg}", flush=True)
    assert rc == _abi.MCRT_ERR_NOMEM and "bytes of device scratch" in msg
    # and the next call still works
    cfg, g, seed = W.plane_rayleigh_tau1(2048, compute_cb=True, Ntheta=40, Nphi=16)
    rc, out, st = _static(cfg, g, seed, "mixed")
    assert rc == 0 and st["rays"] == 2048 and abs(sum(out[:5]) - 1.0) < 1e-5
    with pytest.raises(mcrt._abi.MCRTError):
        _abi.check(_static(cfg, g, seed, "fp64_plain")[0])
