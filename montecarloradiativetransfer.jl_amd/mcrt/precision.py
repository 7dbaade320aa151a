"""Mixed-precision error report: the mixed form against the FP64 form on one Philox stream.

Both forms draw the same uniforms for every ray, so their difference is the
rounding of the FP32 transport and bin evaluation (plus the rare ray whose
history a rounding flips), not Monte Carlo noise.  The report states that
difference per output bin, absolutely and in units of the bin's Monte Carlo
standard error σ (batch means of the FP64 solve), for the energy budget A, the
peel-off Stokes components I, Q, U, V, the CB Stokes components and the CB
enhancement ζ(θb) = ⟨Icb_I(θb, ·)⟩_ϕ / ⟨Icb_I(θb_ref, ·)⟩_ϕ, θb_ref the largest θb up to 60°
with a non-zero CB intensity (or a chosen angle).

Large solves run as consecutive ray ranges of the stream (shards), merged by
ray-weighted means; the union is exactly the rays of one solve of that size.
"""
from __future__ import annotations

import math
import time

import numpy as np

from .api import Plan

GROUPS = ["A", "Irt_I", "Irt_Q", "Irt_U", "Irt_V", "Icb_I", "Icb_Q", "Icb_U", "Icb_V",
          "cb_enhancement"]


def solve_sharded(cfg, g, I0, seed, precision, *, chunk=None, device=0, progress=None,
                  ray_offset=0):
    """One solve of cfg.number_of_rays rays (stream positions ray_offset + 0 ..
    N-1) as shards of at most `chunk` rays.

    Returns (A[5], Irt[4,Nθ,Nϕ], Icb[4,Nθb,Nϕb]) and their standard errors, all
    normalised by the total ray count, plus the summed work counters."""
    N = int(cfg.number_of_rays)
    chunk = N if not chunk else int(chunk)
    vals = [0.0, 0.0, 0.0]
    var = [0.0, 0.0, 0.0]
    stats_tot: dict = {}
    r0 = 0
    while r0 < N:
        n = min(chunk, N - r0)
        plan = Plan(cfg.replace(number_of_rays=n), g, I0, seed=seed, precision=precision,
                    device=device, ray_begin=ray_offset + r0)
        try:
            plan.launch()
            st, sg = {}, {}
            out = plan.read(stats=st, sigma=sg)
        finally:
            plan.close()
        w = n / N
        parts = (np.asarray(out[:5]), out[5], out[6])
        sparts = (sg["A"], sg["Irt"], sg["Icb"])
        for i in range(3):
            vals[i] = vals[i] + w * parts[i]
            var[i] = var[i] + (w * sparts[i]) ** 2
        for k, v in st.items():
            stats_tot[k] = stats_tot.get(k, 0) + v
        r0 += n
        if progress:
            progress(f"{precision}: rays {ray_offset + r0} of [{ray_offset}, {ray_offset + N}) "
                     f"kernel {st['kernel_ms']:.1f} ms")
    sig = [np.sqrt(v) for v in var]
    return tuple(vals), tuple(sig), stats_tot


def enhancement(Icb, sIcb, costhetab, ref_index):
    """ζ(θb) = ⟨I(θb)⟩_ϕ / ⟨I(θb_ref)⟩_ϕ and its standard error (first order,
    covariance between θb rows neglected)."""
    m = Icb[0].mean(axis=1)
    s = np.sqrt((sIcb[0] ** 2).sum(axis=1)) / Icb.shape[2]
    ref, sref = m[ref_index], s[ref_index]
    z = m / ref
    sz = np.abs(z) * np.sqrt((s / np.where(m != 0, m, np.inf)) ** 2 + (sref / ref) ** 2)
    return z, sz


def _group_stats(a, b, s):
    """Bins whose σ is below 1e-9 of their value are deterministic (every ray
    adds the same amount, e.g. the direct transmission at normal incidence):
    they enter the absolute and relative differences, not the |Δ|/σ figures."""
    a = np.asarray(a, float).ravel()
    b = np.asarray(b, float).ravel()
    s = np.asarray(s, float).ravel()
    d = np.abs(a - b)
    live = s > 1e-9 * np.abs(b)
    live &= s > 0
    r = np.full(d.shape, np.nan)
    r[live] = d[live] / s[live]
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    out = dict(bins=int(a.size), bins_with_sigma=int(live.sum()),
               nonfinite=int((~np.isfinite(a)).sum() + (~np.isfinite(b)).sum()),
               max_abs_diff=float(d.max()) if d.size else 0.0,
               max_rel_diff=float(d.max() / scale) if scale > 0 else 0.0)
    det = ~live & (np.abs(b) > 0)
    if det.any():
        out["max_rel_diff_deterministic"] = float(np.max(d[det] / np.abs(b[det])))
    if live.any():
        rl = r[live]
        out.update(median_diff_over_sigma=float(np.median(rl)),
                   p99_diff_over_sigma=float(np.quantile(rl, 0.99)),
                   max_diff_over_sigma=float(rl.max()),
                   frac_above_0p1_sigma=float(np.mean(rl > 0.1)),
                   median_sigma=float(np.median(s[live])))
    return out


def compare_solves(cfg, fp64, mixed, *, ref_theta_b_deg=None):
    """Per-group statistics of |mixed - fp64| and |mixed - fp64|/σ_fp64."""
    (A64, Irt64, Icb64), (sA, sIrt, sIcb) = fp64[0], fp64[1]
    (A32, Irt32, Icb32) = mixed[0]
    rep = {"A": _group_stats(A32, A64, sA)}
    for c, name in enumerate("IQUV"):
        rep[f"Irt_{name}"] = _group_stats(Irt32[c], Irt64[c], sIrt[c])
    if cfg.compute_cb and cfg.Nthetab > 1:
        for c, name in enumerate("IQUV"):
            rep[f"Icb_{name}"] = _group_stats(Icb32[c], Icb64[c], sIcb[c])
        thb = np.degrees(np.arccos(np.clip(np.asarray(cfg.costhetab, float), -1, 1)))
        if ref_theta_b_deg is None:
            # the largest θb up to 60° whose CB intensity is non-zero (towards 90°
            # the backscattered rays graze the slab and carry almost nothing)
            mean = np.asarray(Icb64[0]).mean(axis=1)
            ok = np.nonzero((thb <= 60.0 + 1e-9) & (mean > 0))[0]
            ref_index = int(ok[-1]) if ok.size else len(thb) - 1
        else:
            ref_index = int(np.argmin(np.abs(thb - ref_theta_b_deg)))
        z64, sz = enhancement(Icb64, sIcb, cfg.costhetab, ref_index)
        z32, _ = enhancement(Icb32, sIcb, cfg.costhetab, ref_index)
        st = _group_stats(z32, z64, sz)
        st["theta_b_ref_deg"] = float(thb[ref_index])
        st["zeta_at_0"] = float(z64[0])
        rep["cb_enhancement"] = st
    return rep


def precision_report(cfg, g, I0, seed, *, chunk=None, device=0, progress=None,
                     ref_theta_b_deg=None):
    """Run the FP64 and mixed forms on the same stream and compare them."""
    t = {}
    res = {}
    for form in ("fp64", "mixed"):
        t0 = time.perf_counter()
        res[form] = solve_sharded(cfg, g, I0, seed, form, chunk=chunk, device=device,
                                  progress=progress)
        t[form] = time.perf_counter() - t0
    rep = compare_solves(cfg, res["fp64"], res["mixed"], ref_theta_b_deg=ref_theta_b_deg)
    N = int(cfg.number_of_rays)
    return dict(rays=N, seed=seed, I0=list(map(float, I0)), chunk=chunk or N,
                seconds={k: v for k, v in t.items()},
                rays_per_s={k: N / v for k, v in t.items()},
                events=int(res["fp64"][2].get("events", 0)),
                events_mixed=int(res["mixed"][2].get("events", 0)),
                groups=rep)


def format_report(r) -> str:
    """A markdown table of a precision_report result."""
    lines = [f"rays {r['rays']:.3g} (seed {r['seed']}, shards of {r['chunk']:.3g}); events fp64 "
             f"{r['events']} / mixed {r['events_mixed']}",
             "",
             "| output | bins | max |Δ| | max |Δ|/max|x| | median |Δ|/σ | p99 |Δ|/σ | max |Δ|/σ | "
             "frac |Δ|>0.1σ |",
             "|---|---|---|---|---|---|---|---|"]
    for name in GROUPS:
        s = r["groups"].get(name)
        if not s:
            continue
        f = lambda k: (f"{s[k]:.3g}" if k in s and not math.isnan(s[k]) else "—")
        lines.append(f"| {name} | {s['bins']} | {f('max_abs_diff')} | {f('max_rel_diff')} | "
                     f"{f('median_diff_over_sigma')} | {f('p99_diff_over_sigma')} | "
                     f"{f('max_diff_over_sigma')} | {f('frac_above_0p1_sigma')} |")
    return "\n".join(lines)


def save_solve(path, cfg, res, *, ray_begin, rays, form, seed):
    """Write one sharded solve (values, σ, counters) to an .npz file."""
    (A, Irt, Icb), (sA, sIrt, sIcb), st = res
    np.savez(path, A=A, Irt=Irt, Icb=Icb, sA=sA, sIrt=sIrt, sIcb=sIcb, ray_begin=ray_begin,
             rays=rays, form=form, seed=seed, events=st.get("events", 0),
             kernel_ms=st.get("kernel_ms", 0.0))


def load_merged(paths):
    """Merge saved solves over disjoint ray ranges of one stream (ray-weighted)."""
    parts = [np.load(p, allow_pickle=False) for p in paths]
    N = sum(int(z["rays"]) for z in parts)
    vals = [0.0, 0.0, 0.0]
    var = [0.0, 0.0, 0.0]
    ev = 0
    for z in parts:
        w = int(z["rays"]) / N
        for i, (k, sk) in enumerate((("A", "sA"), ("Irt", "sIrt"), ("Icb", "sIcb"))):
            vals[i] = vals[i] + w * z[k]
            var[i] = var[i] + (w * z[sk]) ** 2
        ev += int(z["events"])
    ranges = sorted((int(z["ray_begin"]), int(z["rays"])) for z in parts)
    for (b0, n0), (b1, _) in zip(ranges, ranges[1:]):
        if b0 + n0 != b1:
            raise ValueError(f"ray ranges are not contiguous: {ranges}")
    return (tuple(vals), tuple(np.sqrt(v) for v in var), {"events": ev}), N, ranges


def main(argv=None) -> int:
    """python -m mcrt.precision [--workload W] [--rays N] [--chunk C]
           python -m mcrt.precision run --form F --rays N --ray-begin R0 --out file.npz
           python -m mcrt.precision report --fp64 a.npz ... --mixed b.npz ...

    The first form runs both precisions and prints the report; `run` stores one
    form over one ray range of the stream, `report` merges stored ranges and
    compares them.  Progress goes to stderr; the report is one JSON line and a
    markdown table."""
    import argparse
    import json
    import sys

    from . import workloads as W

    argv = list(sys.argv[1:] if argv is None else argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", nargs="?", default="both", choices=["both", "run", "report"])
    ap.add_argument("--workload", default="plane_rayleigh_cb", choices=sorted(W.WORKLOADS))
    ap.add_argument("--rays", type=float, default=1e8)
    ap.add_argument("--ray-begin", type=float, default=0)
    ap.add_argument("--chunk", type=float, default=5e7)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--form", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--out", default=None)
    ap.add_argument("--fp64", nargs="*", default=[])
    ap.add_argument("--mixed", nargs="*", default=[])
    a = ap.parse_args(argv)
    I0 = [1.0, -1.0, 0.0, 0.0]
    if a.mode == "run":
        cfg, g, seed = W.WORKLOADS[a.workload](int(a.rays))
        r0 = int(a.ray_begin)
        N = int(cfg.number_of_rays)
        # shards of the range [r0, r0 + N): solve_sharded numbers rays from 0
        vals, var, st = [0.0] * 3, [0.0] * 3, {}
        done = 0
        while done < N:
            n = min(int(a.chunk), N - done)
            res = solve_sharded(cfg.replace(number_of_rays=n), g, I0, seed, a.form,
                                ray_offset=r0 + done, device=a.device, progress=log)
            for i in range(3):
                vals[i] = vals[i] + (n / N) * res[0][i]
                var[i] = var[i] + ((n / N) * res[1][i]) ** 2
            for k, v in res[2].items():
                st[k] = st.get(k, 0) + v
            done += n
        save_solve(a.out, cfg, (tuple(vals), tuple(np.sqrt(v) for v in var), st), ray_begin=r0,
                   rays=N, form=a.form, seed=seed)
        log(f"saved {a.out}: {a.form} rays [{r0}, {r0 + N})")
        return 0
    if a.mode == "report":
        f64, n64, rg64 = load_merged(a.fp64)
        f32, n32, rg32 = load_merged(a.mixed)
        if rg64[0][0] != rg32[0][0] or n64 != n32:
            raise SystemExit(f"fp64 rays {rg64} and mixed rays {rg32} differ")
        cfg, g, seed = W.WORKLOADS[a.workload](n64)
        rep = dict(rays=n64, seed=seed, I0=I0, chunk=a.chunk, events=f64[2]["events"],
                   events_mixed=f32[2]["events"], ranges=rg64,
                   groups=compare_solves(cfg, f64, f32), workload=a.workload)
        print(json.dumps(rep), flush=True)
        print(format_report(rep), flush=True)
        return 0
    cfg, g, seed = W.WORKLOADS[a.workload](int(a.rays))
    rep = precision_report(cfg, g, I0, seed, chunk=int(a.chunk), device=a.device, progress=log)
    rep["workload"] = a.workload
    print(json.dumps(rep), flush=True)
    print(format_report(rep), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
