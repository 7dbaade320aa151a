"""Autotuner for the launch shape of a cooperative solve.

The shape (ray slots per 512-thread workgroup, workgroups launched per CU)
changes how many events share an iteration, how the per-iteration costs (slab
flush, barriers, the transport phase) are amortised and how the tail of each
workgroup's ray range is balanced.  `autotune` times plans of each candidate
shape on a sample of the solve's own rays (a ray range of its stream) and
keeps the fastest per solve kind and size class.  Results do not depend on the
shape beyond FP rounding of the per-iteration partial sums.
"""
from __future__ import annotations

import math
import time

from .api import Plan

_CACHE: dict = {}


def _key(cfg, g, precision):
    kind = type(g).__name__
    sca = type(getattr(g, "scatterer", None) or g.layers[0].scatterer).__name__
    bins = (cfg.Ntheta, cfg.Nphi, cfg.Nthetab if cfg.compute_cb else 0, cfg.Nphib if cfg.compute_cb else 0)
    return (kind, sca, bool(cfg.compute_cb), precision, bins, int(math.log2(max(cfg.number_of_rays, 1))))


def candidates(precision: str):
    slots = (64, 128, 192, 256) if precision == "mixed" else (64, 128)
    return [(s, w) for s in slots for w in (1, 2)]


def autotune(cfg, g, I0, *, precision: str = "mixed", seed: int = 1, sample_rays: int | None = None,
             device: int = 0, cache: dict | None = None, log=None) -> dict:
    """The fastest launch shape for this solve.  Returns {"shape": (slots,
    workgroups_per_cu), "rays_per_s": …, "trials": [...]}; cached per solve kind
    and power-of-two size class."""
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
    for shape in candidates(precision):
        p = Plan(sub, g, I0, seed=seed, precision=precision, device=device, shape=shape,
                 ray_begin=N)  # rays past the solve's own range: a fresh sample of the stream
        t0 = time.perf_counter()
        p.launch()
        p.wait()
        wall = time.perf_counter() - t0
        ms = p.kernel_ms()
        info = p.shape()
        p.close()
        trials.append({"slots": shape[0], "workgroups_per_cu": shape[1],
                       "workgroups": info["workgroups"], "kernel_ms": ms, "wall_s": wall,
                       "rays_per_s": n / (ms * 1e-3)})
        if log:
            log(f"tune {shape}: {trials[-1]['rays_per_s']:.4g} rays/s ({ms:.1f} ms, "
                f"{info['workgroups']} workgroups)")
    best = max(trials, key=lambda t: t["rays_per_s"])
    res = {"shape": (best["slots"], best["workgroups_per_cu"]), "rays_per_s": best["rays_per_s"],
           "sample_rays": n, "trials": trials}
    cache[key] = res
    return res
