"""Autotuner for the launch shape of a cooperative solve.

The shape has three axes:

  * ray slots per 512-thread workgroup (rays in flight; slots/512 rays per
    thread of the transport phase): how many events share an iteration, so how
    the per-iteration costs (slab flush, barriers, the transport and refill
    phases) are amortised, how the tail of each ray range is balanced, and how
    much LDS the event records take;
  * workgroups launched per CU (plans only: a batched or served launch runs
    one workgroup per CU);
  * the LDS table layout: the mixed form with CB reads the amplitude table S
    from LDS (staged beside F and csrn) or through the caches from HBM, which
    frees LDS for more slots when the tables are large.

`autotune` times plans of each candidate shape on a sample of the solve's own
stream; `autotune_batch` times the solve's chunks in a persistent batched
launch, as a Batch or a Server runs them.  `apply_tuned` writes the chosen
(slots, stage_s) into batch or server requests.  Results do not depend on the
shape beyond FP rounding of the per-iteration partial sums.
"""
from __future__ import annotations

import copy
import math
import time

from .api import Plan

_CACHE: dict = {}
_BATCH_CACHE: dict = {}


def _sca_key(g):
    s = getattr(g, "scatterer", None)
    if s is None:
        ss = [L.scatterer for L in g.layers]
        return tuple((type(x).__name__, getattr(x, "Ntheta", 0)) for x in ss)
    return ((type(s).__name__, getattr(s, "Ntheta", 0)),)


def _key(cfg, g, precision, size=True):
    kind = type(g).__name__
    bins = (cfg.Ntheta, cfg.Nphi, cfg.Nthetab if cfg.compute_cb else 0, cfg.Nphib if cfg.compute_cb else 0)
    k = (kind, _sca_key(g), bool(cfg.compute_cb), precision, bins)
    return k + ((int(math.log2(max(cfg.number_of_rays, 1))),) if size else ())


def _tabulated(g) -> bool:
    from .scatterers import RayleighScatterer
    s = getattr(g, "scatterer", None)
    ss = [s] if s is not None else [L.scatterer for L in g.layers]
    return not isinstance(ss[0], RayleighScatterer)


def stage_choices(cfg, g, precision: str):
    """S placements worth trying: both when the mixed form stages S (CB on, a
    tabulated scatterer in layer 1), else the default only."""
    if precision == "mixed" and cfg.compute_cb and _tabulated(g):
        return (1, -1)
    return (0,)


def candidates(precision: str, stages=(0,)):
    slots = (64, 128, 192, 256) if precision == "mixed" else (64, 128)
    return [(s, w, st) for s in slots for w in (1, 2) for st in stages]


def autotune(cfg, g, I0, *, precision: str = "mixed", seed: int = 1, sample_rays: int | None = None,
             device: int = 0, cache: dict | None = None, log=None) -> dict:
    """The fastest plan launch shape for this solve.  Returns {"shape": (slots,
    workgroups_per_cu, stage_s), "rays_per_s": …, "trials": [...]}; cached per
    solve kind and power-of-two size class.  A candidate whose LDS does not fit
    is skipped."""
    import mcrt
    cache = _CACHE if cache is None else cache
    key = _key(cfg, g, precision)
    if key in cache:
        return cache[key]
    N = int(cfg.number_of_rays)
    n = min(N, sample_rays or 1 << 19)
    sub = cfg.replace(number_of_rays=n)
    trials = []
    # one untimed launch warms the code objects and the device
    warm = Plan(sub.replace(number_of_rays=min(n, 4096)), g, I0, seed=seed, precision=precision,
                device=device)
    warm.launch()
    warm.wait()
    warm.close()
    for shape in candidates(precision, stage_choices(cfg, g, precision)):
        try:
            p = Plan(sub, g, I0, seed=seed, precision=precision, device=device, shape=shape,
                     ray_begin=N)  # rays past the solve's own range: a fresh sample of the stream
        except mcrt._abi.MCRTError as e:
            if e.code == mcrt._abi.MCRT_ERR_UNSUPPORTED:
                continue
            raise
        t0 = time.perf_counter()
        p.launch()
        p.wait()
        wall = time.perf_counter() - t0
        ms = p.kernel_ms()
        info = p.shape()
        p.close()
        trials.append({"slots": info["slots"], "workgroups_per_cu": info["workgroups_per_cu"],
                       "stage_s": info["stage_s"], "rays_per_thread": info["slots"] / 512,
                       "workgroups": info["workgroups"], "kernel_ms": ms, "wall_s": wall,
                       "rays_per_s": n / (ms * 1e-3)})
        if log:
            log(f"tune {shape}: {trials[-1]['rays_per_s']:.4g} rays/s ({ms:.1f} ms, "
                f"{info['workgroups']} workgroups)")
    best = max(trials, key=lambda t: t["rays_per_s"])
    res = {"shape": (best["slots"], best["workgroups_per_cu"], best["stage_s"]),
           "rays_per_s": best["rays_per_s"], "sample_rays": n, "trials": trials}
    cache[key] = res
    return res


def autotune_batch(req, *, sample_chunks: int = 256, device: int = 0, cache: dict | None = None,
                   log=None) -> dict:
    """The fastest (slots, stage_s) for the chunks of a batched or served solve:
    each candidate runs `sample_chunks` chunks of the request (a later range of
    its stream) in one persistent batched launch on every CU.  Cached per solve
    kind (geometry, scatterer types and table sizes, CB, precision, bins).
    Returns {"slots": …, "stage_s": …, "rays_per_s": …, "trials": [...]}."""
    import mcrt
    from .batch import Batch
    cache = _BATCH_CACHE if cache is None else cache
    key = _key(req.cfg, req.g, req.precision, size=False)
    if key in cache:
        return cache[key]
    chunk = req.chunk_rays or 2048
    n = int(min(max(req.cfg.number_of_rays, 2 * chunk), sample_chunks * chunk))
    base = copy.copy(req)
    base.cfg = req.cfg.replace(number_of_rays=n)
    base.chunk_rays = chunk
    base.ray_begin = int(req.ray_begin) + int(req.cfg.number_of_rays)  # a fresh sample
    trials = []
    stages = stage_choices(req.cfg, req.g, req.precision)
    slots_axis = (64, 128, 192, 256) if req.precision == "mixed" else (64, 128)
    for slots in slots_axis:
        for st in stages:
            r = copy.copy(base)
            r.slots, r.stage_s = slots, st
            try:
                with Batch([r], device=device) as b:
                    b.launch()
                    b.wait()
                    ms = b.kernel_ms()
            except mcrt._abi.MCRTError as e:
                if e.code == mcrt._abi.MCRT_ERR_UNSUPPORTED:
                    continue
                raise
            trials.append({"slots": slots, "stage_s": st, "rays_per_thread": slots / 512,
                           "kernel_ms": ms, "rays_per_s": n / (ms * 1e-3)})
            if log:
                log(f"tune batch {key[:2]} slots={slots} stage_s={st}: "
                    f"{trials[-1]['rays_per_s']:.4g} rays/s")
    best = max(trials, key=lambda t: t["rays_per_s"])
    res = {"slots": best["slots"], "stage_s": best["stage_s"], "rays_per_s": best["rays_per_s"],
           "sample_rays": n, "trials": trials}
    cache[key] = res
    return res


def apply_tuned(reqs, *, sample_chunks: int = 256, device: int = 0, cache: dict | None = None,
                log=None):
    """Set each request's (slots, stage_s) to the tuned shape of its solve kind
    (autotune_batch, once per kind).  Returns the distinct tuning results."""
    seen = {}
    for r in reqs:
        res = autotune_batch(r, sample_chunks=sample_chunks, device=device, cache=cache, log=log)
        r.slots, r.stage_s = res["slots"], res["stage_s"]
        seen[_key(r.cfg, r.g, r.precision, size=False)] = res
    return seen
