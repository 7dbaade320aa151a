"""simulate and run_rt — src/MonteCarloRadiativeTransfer.jl:26-291.

``simulate(cfg, g, I0)`` is the drop-in boundary: it hands the solve to
libmcrt_hip.so (``mcrt_simulate``) and returns the reference's 7-tuple
``(Adt, Aabs, Aref, Atrans, Arem, Irt, Icb)`` normalised by N, with Irt and
Icb as (4, Nθ, Nϕ) and (4, Nθb, Nϕb) arrays indexed like Julia's.  There is
no CPU fallback: without the library or a GPU the call raises.

Beyond the reference's arguments, ``seed`` keys the Philox stream and
``precision`` picks the arithmetic form ("fp64_plain", "fp64", "mixed").
"""
from __future__ import annotations

import ctypes as C
import math
import threading

import numpy as np

from . import _abi
from .config import Config
from .geometries import AbstractGeometry, PlaneGeometry, SphereGeometry
from .geometries import to_struct as geom_struct

DEFAULT_SEED = 42


class Context:
    """One HIP device context of the library (stream, scratch)."""

    def __init__(self, device: int = 0):
        self.lib = _abi.load_library()
        h = C.c_void_p()
        _abi.check(self.lib.mcrt_context_create(int(device), C.byref(h)))
        self.handle = h
        self.device = device

    def close(self) -> None:
        if self.handle:
            self.lib.mcrt_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ctx_lock = threading.Lock()
_contexts: dict[int, Context] = {}


def get_context(device: int = 0) -> Context:
    with _ctx_lock:
        if device not in _contexts:
            _contexts[device] = Context(device)
        return _contexts[device]


def marshal(cfg: Config, g: AbstractGeometry, I0):
    """ctypes structs for (cfg, g, I0) plus the objects that keep them alive."""
    keep: list = []
    cs = cfg.to_struct(keep)
    gs = geom_struct(g, keep)
    i0 = _abi.f64(I0)
    if i0.shape != (4,):
        raise ValueError("I0 must be a Stokes 4-vector")
    keep.append(i0)
    return cs, gs, i0, keep


def _shape_outputs(cfg: Config, A, Irt, Icb):
    Irt = Irt.reshape(cfg.Nphi, cfg.Ntheta, 4).transpose(2, 1, 0)
    Icb = Icb.reshape(cfg.Nphib, cfg.Nthetab, 4).transpose(2, 1, 0)
    return (A[0], A[1], A[2], A[3], A[4], Irt, Icb)


def simulate(cfg: Config, g: AbstractGeometry, I0, *, use_gpu: bool = True, seed: int = DEFAULT_SEED,
             precision: str = "fp64", device: int = 0, stats: dict | None = None,
             sigma: dict | None = None, ray_begin: int = 0, instance: int = 0):
    """simulate(cfg, g, I₀; use_gpu) — MonteCarloRadiativeTransfer.jl:26-85.

    Returns the reference's 7-tuple (Adt, Aabs, Aref, Atrans, Arem, Irt, Icb), all
    divided by N.  ``stats`` (dict) receives the exact work counters; ``sigma``
    (dict) receives the Monte Carlo standard error of every output ("A", "Irt",
    "Icb", shaped like the outputs) and the number of batches behind it
    (cooperative forms only).  ``ray_begin``/``instance`` select the rays
    ray_begin .. ray_begin+N-1 of the Philox stream (seed, instance)."""
    if not use_gpu:
        raise NotImplementedError(
            "use_gpu=false is the reference's Julia CPU path; this package provides the GPU path")
    if sigma is not None or ray_begin or instance:
        plan = Plan(cfg, g, I0, seed=seed, precision=precision, device=device,
                    ray_begin=ray_begin, instance=instance)
        try:
            plan.launch()
            return plan.read(stats=stats, sigma=sigma)
        finally:
            plan.close()
    g.validate()
    ctx = get_context(device)
    cs, gs, i0, keep = marshal(cfg, g, I0)
    A = np.zeros(5)
    Irt = np.zeros(4 * cfg.Ntheta * cfg.Nphi)
    Icb = np.zeros(max(1, 4 * cfg.Nthetab * cfg.Nphib))
    st = _abi.mcrt_stats()
    rc = ctx.lib.mcrt_simulate(ctx.handle, C.byref(cs), C.byref(gs), _abi.dptr(i0),
                               C.c_uint64(seed & (2**64 - 1)), _abi.PRECISIONS[precision],
                               _abi.dptr(A), _abi.dptr(Irt), _abi.dptr(Icb), C.byref(st))
    _abi.check(rc, ctx.handle)
    if stats is not None:
        stats.update(st.as_dict())
    return _shape_outputs(cfg, A, Irt, Icb[: 4 * cfg.Nthetab * cfg.Nphib])


class Plan:
    """A resident solve: inputs staged in HBM once, launched any number of times."""

    def __init__(self, cfg: Config, g: AbstractGeometry, I0, *, seed: int = DEFAULT_SEED,
                 precision: str = "fp64", device: int = 0, ray_begin: int = 0, instance: int = 0,
                 shape: tuple | None = None):
        g.validate()
        self.cfg = cfg
        self.ctx = get_context(device)
        cs, gs, i0, keep = marshal(cfg, g, I0)
        h = C.c_void_p()
        if ray_begin < 0 or not 0 <= instance < 2**32:
            raise ValueError("ray_begin must be >= 0 and instance a 32-bit unsigned integer")
        sh = _abi.mcrt_launch_shape(*(tuple(shape or ()) + (0, 0, 0))[:3])
        rc = self.ctx.lib.mcrt_plan_create_shaped(self.ctx.handle, C.byref(cs), C.byref(gs),
                                                  _abi.dptr(i0), C.c_uint64(seed & (2**64 - 1)),
                                                  _abi.PRECISIONS[precision], C.c_uint64(ray_begin),
                                                  C.c_uint32(instance), C.byref(sh), C.byref(h))
        _abi.check(rc, self.ctx.handle)
        self.handle = h

    def shape(self) -> dict:
        """Launch shape: slots per workgroup, workgroups per CU, where S is
        read (1 LDS, -1 HBM), grid size."""
        sh = _abi.mcrt_launch_shape()
        wg = C.c_int32()
        _abi.check(self.ctx.lib.mcrt_plan_shape(self.handle, C.byref(sh), C.byref(wg)),
                   self.ctx.handle)
        return {"slots": sh.slots, "workgroups_per_cu": sh.workgroups_per_cu,
                "stage_s": sh.stage_s, "workgroups": wg.value}

    def launch(self) -> None:
        _abi.check(self.ctx.lib.mcrt_plan_launch(self.handle), self.ctx.handle)

    def eval(self, what: str, records) -> np.ndarray:
        """mcrt_plan_eval: the plan's device building block `what` (a key of
        _abi.EVAL) on records of up to 16 inputs each; returns (n, 4)."""
        rec = np.atleast_2d(np.asarray(records, dtype=np.float64))
        inp = np.zeros((rec.shape[0], 16))
        inp[:, :rec.shape[1]] = rec
        out = np.zeros((rec.shape[0], 4))
        _abi.check(self.ctx.lib.mcrt_plan_eval(self.handle, _abi.EVAL[what], rec.shape[0],
                                               _abi.dptr(inp), _abi.dptr(out)), self.ctx.handle)
        return out

    def wait(self) -> None:
        _abi.check(self.ctx.lib.mcrt_plan_wait(self.handle), self.ctx.handle)

    def kernel_ms(self) -> float:
        ms = C.c_double()
        _abi.check(self.ctx.lib.mcrt_plan_kernel_ms(self.handle, C.byref(ms)), self.ctx.handle)
        return ms.value

    def read(self, stats: dict | None = None, sigma: dict | None = None):
        cfg = self.cfg
        A = np.zeros(5)
        Irt = np.zeros(4 * cfg.Ntheta * cfg.Nphi)
        Icb = np.zeros(max(1, 4 * cfg.Nthetab * cfg.Nphib))
        st = _abi.mcrt_stats()
        _abi.check(self.ctx.lib.mcrt_plan_read(self.handle, _abi.dptr(A), _abi.dptr(Irt),
                                               _abi.dptr(Icb), C.byref(st)), self.ctx.handle)
        if stats is not None:
            stats.update(st.as_dict())
        if sigma is not None:
            sA = np.zeros(5)
            sIrt = np.zeros_like(Irt)
            sIcb = np.zeros_like(Icb)
            nb = C.c_int32()
            _abi.check(self.ctx.lib.mcrt_plan_read_sigma(self.handle, _abi.dptr(sA), _abi.dptr(sIrt),
                                                         _abi.dptr(sIcb), C.byref(nb)),
                       self.ctx.handle)
            out = _shape_outputs(cfg, sA, sIrt, sIcb[: 4 * cfg.Nthetab * cfg.Nphib])
            sigma.update({"A": np.asarray(out[:5]), "Irt": out[5], "Icb": out[6],
                          "batches": nb.value})
        return _shape_outputs(cfg, A, Irt, Icb[: 4 * cfg.Nthetab * cfg.Nphib])

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.ctx.lib.mcrt_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ run_rt
def mueller_sphere(cfg: Config, runs):
    """Mueller assembly of run_rt(::SphereGeometry), :87-184, from simulate outputs.

    runs: list of (I0, simulate-tuple) for the columns of initial_I(s).
    Returns (MRT[4,4,Nθ], MCB[4,4,Nθb], energy dict).  The azimuth columns are
    NFI[i] = i·Nϕ÷8 (0-based) for any Nϕ, as the reference's i·Nϕ÷8+1 (:94-95):
    the assembly is exact when Nϕ is a multiple of 8 and uses the nearest lower
    columns otherwise."""
    MRT = np.zeros((4, 4, cfg.Ntheta))
    MCB = np.zeros((4, 4, cfg.Nthetab))
    NFI = [i * cfg.Nphi // 8 for i in range(8)]
    NFIB = [i * cfg.Nphib // 8 for i in range(8)]
    G = dict(dt=0.0, abs=0.0, ref=0.0, rem=0.0)
    for I, (Adt, Aabs, Aref, _At, Arem, IRT, ICB) in runs:
        if I[3] == 0:  # linear
            G["dt"] += Adt; G["abs"] += Aabs; G["ref"] += Aref; G["rem"] += Arem
            for M, X, F in ((MRT, IRT, NFI), (MCB, ICB, NFIB)):
                M[:, 0, :] += (X[:, :, F[0]] + X[:, :, F[2]] + X[:, :, F[4]] + X[:, :, F[6]]) * 0.25
                M[:, 1, :] += (-X[:, :, F[0]] + X[:, :, F[2]] - X[:, :, F[4]] + X[:, :, F[6]]) * 0.25
                M[:, 2, :] += (X[:, :, F[1]] - X[:, :, F[3]] + X[:, :, F[5]] - X[:, :, F[7]]) * 0.25
        else:
            for M, X, F, n in ((MRT, IRT, NFI, cfg.Nphi), (MCB, ICB, NFIB, cfg.Nphib)):
                M[2, 3, :] -= (X[2, :, F[0]] + X[2, :, F[4]]) * 0.5
                M[3, 3, :] -= X[3, :, :].sum(axis=1) / n
    renorm = 1.0 / (G["dt"] + G["ref"])
    MRT *= renorm / cfg.inorm[None, None, :]
    MCB *= renorm
    return MRT, MCB, G


def mueller_plane(cfg: Config, runs):
    """Mueller assembly of run_rt(::PlaneGeometry), :186-291.

    Returns (MRT[4,4,Nθ,Nϕ], MCB[4,4,Nθb,Nϕb], energy dict)."""
    MRT = np.zeros((4, 4, cfg.Ntheta, cfg.Nphi))
    MCB = np.zeros((4, 4, cfg.Nthetab, cfg.Nphib))
    G = dict(dt=0.0, abs=0.0, ref=0.0, trans=0.0, rem=0.0)
    eps = np.finfo(float).eps
    for I, (Adt, Aabs, Aref, Atrans, Arem, IRT, ICB) in runs:
        k1, k2 = 0, 0.0
        for j1 in range(1, 4):
            if abs(I[j1]) > 1.0 - eps:
                k1, k2 = j1, I[j1]
                break
        if k1 == 1:
            G["dt"] += Adt / 2; G["abs"] += Aabs / 2; G["ref"] += Aref / 2
            G["trans"] += Atrans / 2; G["rem"] += Arem / 2
            MRT[:, 0] += IRT / 2.0
            MCB[:, 0] += ICB / 2.0
        MRT[:, k1] += k2 * IRT / 2.0
        MCB[:, k1] += k2 * ICB / 2.0
    norm = np.abs(1.0 / (4.0 * cfg.costheta * cfg.inorm))
    MRT *= norm[None, None, :, None]
    norm1 = 1.0 / np.abs(4.0 * cfg.costhetab)
    MCB *= norm1[None, None, :, None]
    return MRT, MCB, G


def _frame(columns: dict):
    try:
        import pandas as pd
        return pd.DataFrame(columns)
    except ImportError:  # pragma: no cover
        return columns


class ResidentTables:
    """A scatterer's tables pinned in device memory (mcrt_table_upload): solves
    on the device whose scatterer has the same tables use this copy.  Release
    with release() or a with-block."""

    def __init__(self, scatterer, device: int = 0):
        from .scatterers import to_struct as sca_struct
        self.ctx = get_context(device)
        keep: list = []
        st = sca_struct(scatterer, keep)
        h = C.c_uint64()
        _abi.check(self.ctx.lib.mcrt_table_upload(self.ctx.handle, C.byref(st), C.byref(h)),
                   self.ctx.handle)
        self.handle = h.value

    def release(self) -> None:
        if self.handle:
            _abi.check(self.ctx.lib.mcrt_table_release(self.ctx.handle, C.c_uint64(self.handle)),
                       self.ctx.handle)
            self.handle = 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.release()


def table_cache_stats(device: int = 0) -> dict:
    """Bytes, hits and misses of the device context's scatterer-table cache."""
    ctx = get_context(device)
    b, h, m = C.c_int64(), C.c_uint64(), C.c_uint64()
    _abi.check(ctx.lib.mcrt_table_cache_stats(ctx.handle, C.byref(b), C.byref(h), C.byref(m)),
               ctx.handle)
    return {"bytes": b.value, "hits": h.value, "misses": m.value}


def _scatterers(g: AbstractGeometry) -> list:
    if isinstance(g, SphereGeometry):
        return [g.scatterer]
    return [layer.scatterer for layer in g.layers]


def run_rt(cfg: Config, g: AbstractGeometry, *, use_gpu: bool = True, seed: int = DEFAULT_SEED,
           precision: str = "fp64", device: int = 0, verbose: bool = False):
    """run_rt(cfg, g; use_gpu) — MonteCarloRadiativeTransfer.jl:87-291.

    Runs simulate once per incident Stokes column of initial_I(g) and returns
    (SM_rt, SM_cb) as DataFrames with columns θ[, ϕ], s11 … s44.  Column i
    of initial_I is solved with the Philox key seed + i."""
    g.validate()
    IS = g.initial_I()
    runs = []
    # the tables go to the device once for all columns (the reference's adapt
    # before its simulate loop, :97-104 and :194-199)
    pins = [ResidentTables(s, device) for s in _scatterers(g)] if use_gpu else []
    try:
        for i in range(IS.shape[1]):
            I = IS[:, i].copy()
            if verbose:
                print(f"I = {I}")
            runs.append((I, simulate(cfg, g, I, use_gpu=use_gpu, seed=seed + i,
                                     precision=precision, device=device)))
    finally:
        for t in pins:
            t.release()
    deg = 180.0 / math.pi
    if isinstance(g, SphereGeometry):
        MRT, MCB, G = mueller_sphere(cfg, runs)
        rt = {"θ": deg * np.arccos(cfg.costheta)}
        cb = {"θ": deg * np.arccos(cfg.costhetab)}
        for i in range(4):
            for j in range(4):
                rt[f"s{i + 1}{j + 1}"] = MRT[i, j]
                cb[f"s{i + 1}{j + 1}"] = MCB[i, j]
    else:
        MRT, MCB, G = mueller_plane(cfg, runs)
        rt = {"θ": np.tile(deg * np.arccos(cfg.costheta), cfg.Nphi),
              "ϕ": np.repeat(deg * cfg.phi, cfg.Ntheta)}
        cb = {"θ": np.tile(deg * np.arccos(cfg.costhetab), cfg.Nphib),
              "ϕ": np.repeat(deg * cfg.phib, cfg.Nthetab)}
        for i in range(4):
            for j in range(4):
                rt[f"s{i + 1}{j + 1}"] = MRT[i, j].reshape(-1, order="F")
                cb[f"s{i + 1}{j + 1}"] = MCB[i, j].reshape(-1, order="F")
    if verbose:
        for k, v in G.items():
            print(f"{k} (%) = {100 * v:.4f}")
    return _frame(rt), _frame(cb)
