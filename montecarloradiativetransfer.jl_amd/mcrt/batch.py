"""Batched solves: many (cfg, geometry, I₀) problems in one persistent launch.

    reqs = [Request(cfg, g, I0, seed=..., precision="mixed"), ...]
    with Batch(reqs) as b:
        b.launch(); b.wait()
        A, Irt, Icb = b.read(0)[...]

Each solve is cut into chunks of rays by a rule that depends on the solve
alone, so it gets the same result alone and in any batch up to the order of
FP64 additions (see include/mcrt.h, mcrt_batch_create).

``random_instances`` draws the mixed slab/sphere, Rayleigh/Mie population of
the batched benchmark from a seed.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .api import _shape_outputs, get_context, marshal
from .config import Config, midpoint_polar_grid
from .geometries import Layer, PlaneGeometry, SphereGeometry
from .scatterers import MieScatterer, RayleighScatterer


@dataclass
class Request:
    cfg: Config
    g: object
    I0: tuple = (1.0, -1.0, 0.0, 0.0)
    seed: int = 42
    precision: str = "mixed"
    ray_begin: int = 0
    instance: int = 0
    chunk_rays: int = 0
    # launch shape of the solve's chunks: ray slots per workgroup and where the
    # S table is read (mcrt_launch_shape; 0: the defaults, mcrt.tune.shape_for)
    slots: int = 0
    stage_s: int = 0
    meta: dict = field(default_factory=dict)


class Batch:
    """A set of solves staged in device memory, launched together."""

    def __init__(self, reqs: list[Request], device: int = 0):
        if not reqs:
            raise ValueError("a batch needs at least one request")
        self.ctx = get_context(device)
        self.reqs = list(reqs)
        self._keep = []
        arr = (_abi.mcrt_request * len(reqs))()
        for i, r in enumerate(reqs):
            r.g.validate()
            cs, gs, i0, keep = marshal(r.cfg, r.g, r.I0)
            self._keep.append((cs, gs, i0, keep))
            q = arr[i]
            q.cfg = C.pointer(cs)
            q.geom = C.pointer(gs)
            for c in range(4):
                q.I0[c] = float(i0[c])
            q.seed = r.seed & (2**64 - 1)
            q.ray_begin = int(r.ray_begin)
            q.instance = int(r.instance)
            q.precision = _abi.PRECISIONS[r.precision]
            q.chunk_rays = int(r.chunk_rays)
            q.slots = int(r.slots)
            q.stage_s = int(r.stage_s)
        h = C.c_void_p()
        _abi.check(self.ctx.lib.mcrt_batch_create(self.ctx.handle, len(reqs), arr, C.byref(h)),
                   self.ctx.handle)
        self.handle = h
        self._keep = None  # the library copied what it needs

    def launch(self) -> None:
        _abi.check(self.ctx.lib.mcrt_batch_launch(self.handle), self.ctx.handle)

    def wait(self) -> None:
        _abi.check(self.ctx.lib.mcrt_batch_wait(self.handle), self.ctx.handle)

    def kernel_ms(self) -> float:
        ms = C.c_double()
        _abi.check(self.ctx.lib.mcrt_batch_kernel_ms(self.handle, C.byref(ms)), self.ctx.handle)
        return ms.value

    def info(self) -> dict:
        c, b = C.c_int64(), C.c_int64()
        _abi.check(self.ctx.lib.mcrt_batch_info(self.handle, C.byref(c), C.byref(b)),
                   self.ctx.handle)
        return {"chunks": c.value, "device_bytes": b.value}

    def read(self, i: int, stats: dict | None = None, sigma: dict | None = None):
        """The 7-tuple of simulate for solve i of the last launch."""
        cfg = self.reqs[i].cfg
        nrt = 4 * cfg.Ntheta * cfg.Nphi
        ncb = 4 * cfg.Nthetab * cfg.Nphib
        A, Irt, Icb = np.zeros(5), np.zeros(nrt), np.zeros(max(1, ncb))
        sA, sIrt, sIcb = np.zeros(5), np.zeros(nrt), np.zeros(max(1, ncb))
        st = _abi.mcrt_stats()
        _abi.check(self.ctx.lib.mcrt_batch_read(self.handle, i, _abi.dptr(A), _abi.dptr(Irt),
                                                _abi.dptr(Icb), _abi.dptr(sA), _abi.dptr(sIrt),
                                                _abi.dptr(sIcb), C.byref(st)), self.ctx.handle)
        if stats is not None:
            stats.update(st.as_dict())
        if sigma is not None:
            so = _shape_outputs(cfg, sA, sIrt, sIcb[:ncb])
            sigma.update({"A": np.asarray(so[:5]), "Irt": so[5], "Icb": so[6]})
        return _shape_outputs(cfg, A, Irt, Icb[:ncb])

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.ctx.lib.mcrt_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _loguniform(rng, lo, hi):
    return float(math.exp(rng.uniform(math.log(lo), math.log(hi))))


# θb: 0-5° every 0.05° (101) then 6-180° every 1° (175)
BATCH_THETAB_DEG = np.concatenate([np.arange(101) * 0.05, np.arange(6.0, 181.0)])


def random_instances(n: int = 256, *, seed: int = 2024, ray_scale: float = 1.0,
                     Ntheta_table: int = 1801, random_points: int = 1000, precision="mixed",
                     lam: float = 0.55, nu: float = 0.1, min_rays: int = 1000):
    """The batched-benchmark population, drawn from `seed`:

    50/50 plane slab (normal incidence) / sphere; optical thickness or radius
    τ log-uniform in [0.1, 20]; 25% Rayleigh (ω = 1, kl = 1), else Mie spheres
    with size parameter x uniform in [0.5, 20], Re m uniform in [1.3, 1.7] and
    Im m log-uniform in [1e-5, 1e-1] at λ = 0.55 µm, ν = 0.1; rays log-uniform
    in [1e5, 1e8] times ray_scale (at least min_rays); CB on; RT 180 × 1° polar
    midpoint bins × 48 azimuths; CB θb 0-5° every 0.05° plus 6-180° every 1°,
    8 azimuths.  Returns a list of Request (meta holds the drawn parameters)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        plane = bool(rng.random() < 0.5)
        tau = _loguniform(rng, 0.1, 20.0)
        rayleigh = bool(rng.random() < 0.25)
        x = float(rng.uniform(0.5, 20.0))
        mre = float(rng.uniform(1.3, 1.7))
        mim = _loguniform(rng, 1e-5, 1e-1)
        N = max(min_rays, int(round(_loguniform(rng, 1e5, 1e8) * ray_scale)))
        if rayleigh:
            sca = RayleighScatterer(1.0, 1.0)
        else:
            r = x * lam / (2 * math.pi)
            sca = MieScatterer.create(r, complex(mre, mim), lam, nu, Ntheta=Ntheta_table,
                                      random_points=random_points)
        if plane:
            g = PlaneGeometry(math.pi, 0.0, [Layer(d=tau * sca.kl, scatterer=sca, tau=tau,
                                                   omega=sca.omega)])
        else:
            g = SphereGeometry(R=tau * sca.kl, scatterer=sca)
        ct, st, w = midpoint_polar_grid(g, 180, 48)
        cfg = Config(g, number_of_rays=N, maximum_scattering_times=10000, compute_cb=True,
                     Ntheta=180, Nphi=48, costheta=ct, sintheta=st, inorm=w,
                     thetab=np.deg2rad(BATCH_THETAB_DEG), Nphib=8, minimum_intensity=1e-6,
                     tau0=50.0)
        meta = dict(index=i, geometry="plane" if plane else "sphere", tau=tau,
                    scatterer="rayleigh" if rayleigh else "mie", x=None if rayleigh else x,
                    m=None if rayleigh else (mre, mim), rays=N)
        out.append(Request(cfg, g, seed=seed * 1000 + i, precision=precision, meta=meta))
    return out


def main(argv=None) -> int:
    """python -m mcrt.batch [--instances 256] [--ray-scale 0.01] [--seed 2024] [--launches 2]

    The batched benchmark: the seed-generated population in one persistent
    launch per precision (graph-captured; the first launch captures, later ones
    replay).  Prints one JSON line with rays/s over the timed launches."""
    import argparse
    import json
    import sys
    import time

    from . import opcount

    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--ray-scale", type=float, default=0.01)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--launches", type=int, default=2)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--tuned", action="store_true",
                    help="launch shape of each solve kind from the autotuner (mcrt.tune.apply_tuned)")
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    reqs = random_instances(a.instances, seed=a.seed, ray_scale=a.ray_scale, precision=a.precision)
    rays = sum(r.cfg.number_of_rays for r in reqs)
    log(f"population: {len(reqs)} instances, {rays} rays (scale {a.ray_scale}), "
        f"tables built in {time.perf_counter() - t0:.1f} s")
    kinds = {}
    for r in reqs:
        k = f"{r.meta['geometry']}/{r.meta['scatterer']}"
        kinds[k] = kinds.get(k, 0) + 1
    tuned = None
    if a.tuned:
        from .tune import apply_tuned
        t = time.perf_counter()
        res = apply_tuned(reqs, log=log)
        tuned = {"seconds": time.perf_counter() - t,
                 "shapes": [{"kind": f"{k[0]}/{'+'.join(n for n, _ in k[1])}", "slots": v["slots"],
                             "stage_s": v["stage_s"], "rays_per_s": v["rays_per_s"]}
                            for k, v in res.items()]}
        log(f"tuned: {tuned}")
    with Batch(reqs) as b:
        info = b.info()
        log(f"batch: {info}")
        times = []
        for i in range(a.launches):
            t = time.perf_counter()
            b.launch()
            b.wait()
            times.append((time.perf_counter() - t, b.kernel_ms()))
            log(f"launch {i}: wall {times[-1][0]:.2f} s, device {times[-1][1]:.1f} ms")
        tot = {"events": 0, "rt_evals": 0, "cb_evals": 0, "cb_evals_single": 0, "nonfinite": 0}
        flops = 0.0
        for i, r in enumerate(reqs):
            st = {}
            out = b.read(i, stats=st)
            for k in tot:
                tot[k] += st[k]
            flops += opcount.algorithmic_flops(r.g, st)
            if not abs(sum(out[:5]) - 1.0) < 1e-4:
                log(f"instance {i}: energy budget {sum(out[:5])}")
    timed = times[1:] if len(times) > 1 else times
    wall = sum(t for t, _ in timed) / len(timed)
    dev = sum(d for _, d in timed) / len(timed)
    print(json.dumps({"workload": "batched population", "instances": len(reqs), "rays": rays,
                      "ray_scale": a.ray_scale, "population_seed": a.seed,
                      "precision": a.precision, "kinds": kinds, "chunks": info["chunks"],
                      "device_bytes": info["device_bytes"], "wall_s": wall, "device_ms": dev,
                      "rays_per_s": rays / wall, "tflops_algorithmic": flops / (dev * 1e-3) / 1e12,
                      "tuned": tuned, **tot}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
