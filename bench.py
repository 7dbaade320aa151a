"""Benchmark of the per-ray Monte Carlo RT/CB transport on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one full solve (one launch over all rays of the workload, inputs
resident in HBM).  The workload is the sphere of Mie scatterers with coherent
backscattering: optical radius 5, r = 0.5 µm, λ = 0.55 µm, m = 1.5+0.001i,
ν = 0.1, F/S on 1801 angles, RT 180×48 bins, CB 101×8 bins, 1e7 rays, seed 7.
`value` is rays/s of the fastest built form (mixed); the FP64 forms and the
oracle on the host cores are timed beside it on bounded samples of the same
stream (the same rays, so the Monte Carlo error is matched), and a short
Poisson serving run gives the p50/p99 latency.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "montecarloradiativetransfer.jl_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

PEAK_TFLOPS = {"mixed": 157.3, "fp64": 78.6, "fp64_plain": 78.6}  # MI355X vector FP32 / FP64
DTYPE = {"mixed": "f32+f64", "fp64": "f64", "fp64_plain": "f64"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_form(cfg, g, seed, I0, form, steps, warmup):
    import mcrt

    plan = mcrt.Plan(cfg, g, I0, seed=seed, precision=form)
    for _ in range(warmup):
        plan.launch()
        plan.wait()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.launch()
        plan.wait()
        kms.append(plan.kernel_ms())
    t1 = time.perf_counter()
    st = {}
    out = plan.read(stats=st)
    plan.close()
    return dict(wall_s=t1 - t0, kernel_ms=kms, stats=st, out=out)


def cpu_threads() -> int:
    """Host threads for the CPU baseline: the CPUs this process may run on, at
    most OMP_NUM_THREADS when set (the box's CPU share) and at most 16."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, min(n, 16))


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def measured_traffic(workload: str, rays: int, form: str):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summaries (profiles/*.md lines `traffic_bytes_per_launch[workload, N rays,
    form] = B`), or None when no profile covers this workload, size and form."""
    import re

    pat = re.compile(r"^traffic_bytes_per_launch\[(\w+), (\d+) rays, (\w+)\] = ([0-9.eE+]+)\s*$")
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            m = pat.match(line.strip())
            if m and m.group(1) == workload and int(m.group(2)) == rays and m.group(3) == form:
                return float(m.group(4))
    return None


def copy_calibration(mib: int) -> int:
    """hipMemcpy device→device of mib MiB: one copy kernel that reads and writes
    exactly mib MiB, for scaling the FETCH_SIZE/WRITE_SIZE counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_mib": mib}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sphere_mie_cb")
    ap.add_argument("--rays", type=int, default=None, help="override the workload's ray count")
    ap.add_argument("--form", default="auto", choices=["auto", "mixed", "fp64", "fp64_plain"])
    ap.add_argument("--compare", default="fp64,fp64_plain",
                    help="other forms timed beside the headline form")
    ap.add_argument("--plain-sample", type=int, default=65536,
                    help="rays of the plain FP64 comparison run (8 per thread at the "
                         "reference's maximum_threads = 8192, so that it measures throughput "
                         "rather than the slowest single ray)")
    ap.add_argument("--fp64-sample", type=int, default=2_000_000,
                    help="rays of the FP64 comparison run")
    ap.add_argument("--cpu-sample", type=int, default=4096,
                    help="rays of each of the 3 oracle CPU baseline runs (the median is reported)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="seconds of Poisson serving load (200 req/s) for the p50/p99 latency; 0: skip")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (the known-size kernel "
                         "that scales rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(json.dumps({"error": "this build targets one GPU; --gpus must be 1 (rays shard "
                          "across GPUs by disjoint ray ranges with one final bin reduction, "
                          "which is not built here)"}))
        return 2

    import mcrt
    from mcrt import opcount
    from mcrt import workloads as W

    make = W.WORKLOADS[args.workload]
    cfg, g, seed = make() if args.rays is None else make(args.rays)
    I0 = [1.0, -1.0, 0.0, 0.0]
    N = cfg.number_of_rays

    forms = ["mixed", "fp64", "fp64_plain"] if args.form == "auto" else [args.form]
    head = None
    for form in forms:
        try:
            small = cfg.replace(number_of_rays=64)
            mcrt.Plan(small, g, I0, seed=seed, precision=form).close()
            head = form
            break
        except mcrt._abi.MCRTError as e:
            if e.code != mcrt._abi.MCRT_ERR_UNSUPPORTED:
                raise
    if head is None:
        raise SystemExit("no transport form is available")

    log(f"bench: workload={args.workload} N={N} form={head} steps={args.steps} warmup={args.warmup}")
    r = run_form(cfg, g, seed, I0, head, args.steps, args.warmup)
    st = r["stats"]
    value = args.steps * N / r["wall_s"]
    ms_step = 1e3 * r["wall_s"] / args.steps
    kavg = sum(r["kernel_ms"]) / len(r["kernel_ms"])
    flops = opcount.algorithmic_flops(g, st)
    achieved = flops / (kavg * 1e-3) / 1e12
    log(f"bench: {head}: {value:.4g} rays/s, kernel {kavg:.1f} ms, {achieved:.2f} TFLOP/s "
        f"(algorithmic), events/ray {st['events'] / N:.2f}")

    compare = {}
    for form in [f for f in args.compare.split(",") if f and f not in (head, "none")]:
        n = args.plain_sample if form == "fp64_plain" else args.fp64_sample
        n = min(n, N)
        try:
            c = run_form(cfg.replace(number_of_rays=n), g, seed, I0, form, 1, 0)
        except mcrt._abi.MCRTError as e:
            if e.code == mcrt._abi.MCRT_ERR_UNSUPPORTED:
                continue
            raise
        rate = n / c["wall_s"]
        ck = c["kernel_ms"][0]
        cf = opcount.algorithmic_flops(g, c["stats"]) / (ck * 1e-3) / 1e12
        compare[form] = {"rays_per_s": rate, "rays": n, "kernel_ms": ck, "tflops": cf,
                         "speedup_of_headline": value / rate}
        log(f"bench: {form}: {rate:.4g} rays/s on {n} rays ({cf:.2f} TFLOP/s)")

    cpu = None
    if not args.no_cpu:
        import oracle

        nth = cpu_threads()
        ncpu = min(args.cpu_sample, N)
        S = oracle.Solve(cfg.replace(number_of_rays=ncpu), g)
        times = []
        for k in range(3):  # three runs over the same rays, the median reported
            t0 = time.perf_counter()
            S.simulate(I0, seed, nthreads=nth, nbatch=max(nth * 4, 16))
            times.append(time.perf_counter() - t0)
        dt = sorted(times)[1]
        cpu = {"value": ncpu / dt, "unit": "rays/s", "cores": nth, "kind": "port",
               "sample": f"rays 0..{ncpu - 1} of the same workload and seed, median of 3 runs "
                         f"({', '.join(f'{t:.2f}' for t in times)} s; oracle, FP64, {nth} threads "
                         f"of {os.cpu_count()} logical CPUs, {cpu_model()})"}
        log(f"bench: cpu oracle {ncpu / dt:.4g} rays/s on {nth} threads (median of {times})")

    serving = None
    if args.serve_seconds > 0:
        from mcrt.batch import random_instances
        from mcrt.serve import Server, poisson_load

        scale, rate = 5e-5, 200.0
        pool = random_instances(256, seed=2024, ray_scale=scale, min_rays=100)
        with Server() as srv:
            rep = poisson_load(srv, pool, rate=rate, duration=args.serve_seconds, seed=5)
        serving = {"workload": "Poisson arrivals from the 256-instance batched population "
                               "(mixed form, rays scaled by 5e-5, at least 100)",
                   "rate_per_s": rate, "seconds": args.serve_seconds,
                   "requests": rep["requests"], "completed": rep["completed"],
                   "rays_per_s": rep["rays_per_s"], "p50_ms": rep.get("p50_ms"),
                   "p99_ms": rep.get("p99_ms"),
                   # per request: submit → admission (queue), admission → device start of
                   # the first window running one of its chunks (window), then to done (device)
                   "queue_ms": rep["breakdown"]["queue_ms"],
                   "window_wait_ms": rep["breakdown"]["wait_ms"],
                   "device_ms": rep["breakdown"]["device_ms"],
                   "launches": rep["server"]["launches"],
                   "preemptions": rep["server"]["preemptions"],
                   "resumed_chunks": rep["server"]["resumed"],
                   "table_hits": rep["server"]["table_hits"],
                   "table_misses": rep["server"]["table_misses"]}
        log(f"bench: serving {rep['completed']}/{rep['requests']} requests, "
            f"{rep['rays_per_s']:.4g} rays/s, p50 {rep.get('p50_ms', 0):.1f} ms, "
            f"p99 {rep.get('p99_ms', 0):.1f} ms")

    line = {
        "metric": "rays/sec on 1 GPU, mixed vs FP64 plain at matched MC error; p50/p99 latency",
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[head],
        "data": "synthetic",
        "config": {"workload": args.workload, "rays": N, "form": head, "seed": seed,
                   "rt_bins": [cfg.Ntheta, cfg.Nphi], "cb_bins": [cfg.Nthetab, cfg.Nphib],
                   "mie_table": getattr(g, "scatterer", None) and getattr(g.scatterer, "Ntheta", None),
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS[head],
                     "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS[head],
                     "traffic": measured_traffic(args.workload, N, head),
                     "kernel_ms": kavg,
                     "flops_per_launch": flops,
                     "units_per_launch": {"rt_evals": st["rt_evals"], "cb_evals": st["cb_evals"],
                                          "cb_evals_single": st["cb_evals_single"]}},
        "cpu_baseline": cpu,
        "events_per_ray": st["events"] / N,
        "compare": compare,
        "serving": serving,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
