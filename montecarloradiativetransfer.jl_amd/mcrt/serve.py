"""The request loop: submit solves to a server on the device, collect results.

    with Server() as srv:
        t = srv.submit(Request(cfg, g, I0, seed=1, precision="mixed"))
        out = srv.result(t)            # blocks; the 7-tuple of simulate

``poisson_load`` drives a server with Poisson arrivals drawn from a pool of
requests and reports throughput and latency percentiles; ``python -m
mcrt.serve`` runs that load on the batched-benchmark population.
"""
from __future__ import annotations

import ctypes as C
import threading
import time

import numpy as np

from . import _abi
from .api import _shape_outputs, get_context, marshal
from .batch import Request

STATES = _abi.TICKET_STATES


class Server:
    def __init__(self, device: int = 0, device_budget: int = 0, table_cache: int = 0):
        self.ctx = get_context(device)
        h = C.c_void_p()
        _abi.check(self.ctx.lib.mcrt_server_create(self.ctx.handle, int(device_budget),
                                                   int(table_cache), C.byref(h)), self.ctx.handle)
        self.handle = h
        self._reqs: dict[int, Request] = {}
        self._lock = threading.Lock()

    def submit(self, r: Request) -> int:
        r.g.validate()
        cs, gs, i0, keep = marshal(r.cfg, r.g, r.I0)
        q = _abi.mcrt_request()
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
        t = C.c_uint64()
        _abi.check(self.ctx.lib.mcrt_submit(self.handle, C.byref(q), C.byref(t)))
        with self._lock:
            self._reqs[t.value] = r
        return t.value

    def poll(self, ticket: int) -> str:
        s = C.c_int32()
        _abi.check(self.ctx.lib.mcrt_poll(self.handle, C.c_uint64(ticket), C.byref(s)))
        return STATES[s.value]

    def wait(self, ticket: int, timeout_ms: float = -1.0) -> str:
        s = C.c_int32()
        rc = self.ctx.lib.mcrt_wait(self.handle, C.c_uint64(ticket), float(timeout_ms), C.byref(s))
        if rc not in (_abi.MCRT_OK, _abi.MCRT_ERR_NOT_READY):
            _abi.check(rc)
        return STATES[s.value]

    def result(self, ticket: int, stats: dict | None = None, sigma: dict | None = None):
        """Wait for the ticket and return its 7-tuple (raises MCRTError when the
        ticket was cancelled or failed)."""
        self.wait(ticket)
        with self._lock:
            r = self._reqs.pop(ticket)
        cfg = r.cfg
        nrt = 4 * cfg.Ntheta * cfg.Nphi
        ncb = 4 * cfg.Nthetab * cfg.Nphib
        A, Irt, Icb = np.zeros(5), np.zeros(nrt), np.zeros(max(1, ncb))
        sA, sIrt, sIcb = np.zeros(5), np.zeros(nrt), np.zeros(max(1, ncb))
        st = _abi.mcrt_stats()
        lat = C.c_double()
        _abi.check(self.ctx.lib.mcrt_result(self.handle, C.c_uint64(ticket), _abi.dptr(A),
                                            _abi.dptr(Irt), _abi.dptr(Icb), _abi.dptr(sA),
                                            _abi.dptr(sIrt), _abi.dptr(sIcb), C.byref(st),
                                            C.byref(lat)))
        if stats is not None:
            stats.update(st.as_dict())
            stats["latency_ms"] = lat.value
        if sigma is not None:
            so = _shape_outputs(cfg, sA, sIrt, sIcb[:ncb])
            sigma.update({"A": np.asarray(so[:5]), "Irt": so[5], "Icb": so[6]})
        return _shape_outputs(cfg, A, Irt, Icb[:ncb])

    def cancel(self, ticket: int) -> None:
        _abi.check(self.ctx.lib.mcrt_cancel(self.handle, C.c_uint64(ticket)))

    def set_slice(self, ms: float) -> None:
        """How long a launch window keeps the device while other work waits
        (0: stop it as soon as work waits)."""
        _abi.check(self.ctx.lib.mcrt_server_set_slice(self.handle, float(ms)))

    def stats(self) -> dict:
        s = _abi.mcrt_server_stats()
        _abi.check(self.ctx.lib.mcrt_server_stats_get(self.handle, C.byref(s)))
        return s.as_dict()

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.ctx.lib.mcrt_server_destroy(self.handle)
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


def arrivals(rate: float, duration: float, seed: int) -> np.ndarray:
    """Poisson arrival times (s) in [0, duration) at `rate` per second."""
    rng = np.random.default_rng(seed)
    n = int(rate * duration * 1.5) + 16
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    while t[-1] < duration:
        t = np.concatenate([t, t[-1] + np.cumsum(rng.exponential(1.0 / rate, size=n))])
    return t[t < duration]


def latency_summary(lat_ms) -> dict:
    a = np.asarray(lat_ms, float)
    if a.size == 0:
        return {"n": 0}
    return {"n": int(a.size), "p50_ms": float(np.percentile(a, 50)),
            "p99_ms": float(np.percentile(a, 99)), "mean_ms": float(a.mean()),
            "max_ms": float(a.max())}


def _pcts(v) -> dict:
    a = np.asarray(v, float)
    if a.size == 0:
        return {}
    return {"p50": float(np.percentile(a, 50)), "p99": float(np.percentile(a, 99)),
            "mean": float(a.mean())}


def poisson_load(srv: Server, pool: list[Request], *, rate: float, duration: float,
                 seed: int = 5, keep: int = 0, log=None) -> dict:
    """Submit requests drawn uniformly from `pool` at Poisson arrival times;
    collect every result.  Request i gets stream seed `seed·10⁶ + i`.  The first
    `keep` results are returned (with their requests) for accuracy checks."""
    times = arrivals(rate, duration, seed)
    rng = np.random.default_rng(seed + 1)
    picks = rng.integers(0, len(pool), size=len(times))
    tickets: list[tuple[int, Request]] = []
    done = threading.Event()
    lat, rays, kept = [], [], []
    parts = {"queue_ms": [], "wait_ms": [], "device_ms": []}
    errors = []

    def collect():
        i = 0
        while True:
            if i >= len(tickets):
                if done.is_set() and i >= len(tickets):
                    return
                time.sleep(0.0005)
                continue
            t, r = tickets[i]
            st = {}
            try:
                out = srv.result(t, stats=st)
            except _abi.MCRTError as e:
                errors.append(str(e))
                i += 1
                continue
            lat.append(st["latency_ms"])
            rays.append(st["rays"])
            parts["queue_ms"].append(st["queue_ms"])
            parts["wait_ms"].append(st["wait_ms"])
            parts["device_ms"].append(st["kernel_ms"])
            if len(kept) < keep:
                kept.append((r, out))
            i += 1

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    for i, (ta, k) in enumerate(zip(times, picks)):
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        base = pool[int(k)]
        r = Request(base.cfg, base.g, base.I0, seed=seed * 10**6 + i, precision=base.precision,
                    chunk_rays=base.chunk_rays, meta=dict(base.meta, pick=int(k)))
        tickets.append((srv.submit(r), r))
        if log and (i + 1) % 1000 == 0:
            log(f"submitted {i + 1}/{len(times)} at {time.perf_counter() - t0:.1f} s")
    t_submit = time.perf_counter() - t0
    done.set()
    th.join()
    wall = time.perf_counter() - t0
    out = {"requests": len(times), "rate": rate, "duration": duration, "arrival_seed": seed,
           "submit_s": t_submit, "wall_s": wall, "completed": len(lat), "errors": len(errors),
           "rays": int(sum(rays)), "rays_per_s": float(sum(rays) / wall)}
    out.update(latency_summary(lat))
    # per request: submit → admission, admission → device start of the first
    # window that ran one of its chunks, that start → completion
    out["breakdown"] = {k: _pcts(v) for k, v in parts.items()}
    out["server"] = srv.stats()
    out["kept"] = kept
    return out


def main(argv=None) -> int:
    """python -m mcrt.serve [--rate 200] [--duration 60] [--ray-scale 5e-5] [--check 16]

    Poisson load drawn from the 256-instance batched population (mixed form),
    then an FP64 re-run of the first `--check` requests on the same streams
    for the per-request 3σ agreement.  Prints one JSON line."""
    import argparse
    import json
    import sys

    from .batch import Batch, random_instances

    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--ray-scale", type=float, default=5e-5)
    ap.add_argument("--population-seed", type=int, default=2024)
    ap.add_argument("--arrival-seed", type=int, default=5)
    ap.add_argument("--check", type=int, default=16)
    ap.add_argument("--budget-gib", type=float, default=64.0)
    ap.add_argument("--min-rays", type=int, default=100)
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    pool = random_instances(256, seed=a.population_seed, ray_scale=a.ray_scale,
                            min_rays=a.min_rays)
    log(f"pool: {len(pool)} instances, mean rays {np.mean([r.cfg.number_of_rays for r in pool]):.0f}")
    with Server(device_budget=int(a.budget_gib * 2**30)) as srv:
        rep = poisson_load(srv, pool, rate=a.rate, duration=a.duration, seed=a.arrival_seed,
                           keep=a.check, log=log)
    kept = rep.pop("kept")
    # per-request agreement with FP64 on the same stream (batched FP64 re-run)
    from .stats import compare_results
    reqs64 = [Request(r.cfg, r.g, r.I0, seed=r.seed, precision="fp64", chunk_rays=r.chunk_rays)
              for r, _ in kept]
    checks, nbs = [], []
    if reqs64:
        with Batch(reqs64) as b:
            b.launch()
            b.wait()
            for i, (r, out) in enumerate(kept):
                sg, st = {}, {}
                ref = b.read(i, sigma=sg, stats=st)
                nbs.append(int(st["batches"]))
                checks.append(compare_results(out, ref, sg, same_stream=True, batches=nbs[-1]))
    # χ²/dof is a test only where the σ has 5 or more degrees of freedom; below
    # that the |z| > 3 count (Student t) decides
    chi = [c["chi2dof"] for c, nb in zip(checks, nbs) if nb - 1 >= 5]
    rep["fp64_check"] = {"requests": len(checks),
                         "all_within_3sigma": bool(all(c["ok"] for c in checks)),
                         "worst_exceed3_frac": max((c["exceed3_frac"] for c in checks), default=0.0),
                         "batches": nbs,
                         "chi2_tested": len(chi),
                         "max_chi2dof_tested": max(chi, default=None)}
    print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
