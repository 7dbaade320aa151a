"""Batched solves (one persistent launch per precision) against single solves."""
import numpy as np
import pytest

import mcrt
from mcrt import workloads as W
from mcrt.batch import Batch, Request, random_instances
from statcheck import compare, flat

pytestmark = pytest.mark.gpu

I0 = (1.0, -1.0, 0.0, 0.0)


def _mix():
    a = W.plane_rayleigh_tau1(4096, compute_cb=True)
    b = W.sphere_mie_cb(4096, Ntheta_table=1801)
    c = W.sphere_mie_cb(2048, Ntheta_table=1801)
    d = W.plane_rayleigh_tau1(4096)
    return [Request(a[0], a[1], I0, seed=a[2], precision="mixed"),
            Request(b[0], b[1], I0, seed=b[2], precision="mixed"),
            Request(c[0], c[1], I0, seed=c[2] + 1, precision="fp64"),
            Request(d[0], d[1], I0, seed=d[2], precision="fp64")]


def _tol(form, ref):
    return (1e-5 if form == "mixed" else 1e-11) * float(np.max(np.abs(ref)))


def test_batch_matches_single_solves():
    reqs = _mix()
    print(f"batch of {len(reqs)}", flush=True)
    with Batch(reqs) as b:
        b.launch()
        b.wait()
        print(f"batch info {b.info()} {b.kernel_ms():.1f} ms", flush=True)
        for i, r in enumerate(reqs):
            st = {}
            got = flat(b.read(i, stats=st))
            ref = flat(mcrt.simulate(r.cfg, r.g, r.I0, seed=r.seed, precision=r.precision))
            err = float(np.max(np.abs(got - ref)))
            print(f"[batch/{i}/{r.precision}] rays={st['rays']} max|Δ|={err:.3e} "
                  f"tol={_tol(r.precision, ref):.3e}", flush=True)
            assert st["rays"] == r.cfg.number_of_rays and st["nonfinite"] == 0
            assert abs(sum(got[:5]) - 1.0) < (1e-5 if r.precision == "mixed" else 1e-9)
            assert err <= _tol(r.precision, ref)


def test_alone_equals_in_the_mix():
    reqs = _mix()
    alone = Batch([reqs[1]])
    mixed = Batch(reqs)
    for x in (alone, mixed):
        x.launch()
        x.wait()
    a = flat(alone.read(0))
    m = flat(mixed.read(1))
    err = float(np.max(np.abs(a - m)))
    print(f"[alone-vs-mix] max|Δ|={err:.3e}", flush=True)
    assert err <= 1e-12 * float(np.max(np.abs(a)))
    alone.close()
    mixed.close()


def test_graph_replays_agree_with_the_first_launch():
    reqs = _mix()[:3]
    with Batch(reqs) as b:
        outs = []
        for k in range(3):
            b.launch()
            b.wait()
            outs.append([flat(b.read(i)) for i in range(len(reqs))])
            # eager work of the program between replays
            mcrt.simulate(reqs[0].cfg.replace(number_of_rays=256), reqs[0].g, I0, seed=k,
                          precision="mixed")
        for k in (1, 2):
            for i in range(len(reqs)):
                err = float(np.max(np.abs(outs[k][i] - outs[0][i])))
                assert err <= 1e-12 * float(np.max(np.abs(outs[0][i]))), (k, i, err)


def test_batch_sigma_consistent_with_independent_streams():
    cfg, g, seed = W.sphere_mie_cb(65536, Ntheta_table=1801)
    reqs = [Request(cfg, g, I0, seed=seed, precision="mixed"),
            Request(cfg, g, I0, seed=seed + 1, precision="mixed")]
    with Batch(reqs) as b:
        b.launch()
        b.wait()
        sa, sb = {}, {}
        a = flat(b.read(0, sigma=sa))
        c = flat(b.read(1, sigma=sb))
    s = np.sqrt(flat((*sa["A"], sa["Irt"], sa["Icb"])) ** 2 + flat((*sb["A"], sb["Irt"], sb["Icb"])) ** 2)
    nb = min(128, -(-cfg.number_of_rays // 512))
    r = compare(a, c, s, floor=1e-15, dead_rel=1e-6, dof=nb - 1, label="batch-sigma")
    assert r["ok"], r
    assert r["chi2dof"] > 0.7, r


def test_random_population_runs():
    reqs = random_instances(8, seed=7, ray_scale=1e-3)
    with Batch(reqs) as b:
        b.launch()
        b.wait()
        for i, r in enumerate(reqs):
            st = {}
            out = flat(b.read(i, stats=st))
            print(f"[pop/{i}] {r.meta['geometry']} {r.meta['scatterer']} tau={r.meta['tau']:.2f} "
                  f"rays={st['rays']} events={st['events']}", flush=True)
            assert st["rays"] == r.cfg.number_of_rays
            assert np.all(np.isfinite(out))
            assert abs(sum(out[:5]) - 1.0) < 1e-5


def test_launch_shapes_and_autotuner():
    from mcrt.api import Plan
    from mcrt.tune import autotune
    cfg, g, seed = W.sphere_mie_cb(32768, Ntheta_table=1801)
    ref = flat(mcrt.simulate(cfg, g, I0, seed=seed, precision="mixed"))
    for shape in [(64, 1), (128, 2), (256, 1)]:
        p = Plan(cfg, g, I0, seed=seed, precision="mixed", shape=shape)
        info = p.shape()
        assert info["slots"] == shape[0] and info["workgroups_per_cu"] == shape[1]
        p.launch()
        out = flat(p.read())
        p.close()
        err = float(np.max(np.abs(out - ref)))
        print(f"[shape {shape}] workgroups={info['workgroups']} max|Δ|={err:.3e}", flush=True)
        assert err <= 1e-5 * float(np.max(np.abs(ref)))  # same rays, FP32 grouping only
    res = autotune(cfg, g, I0, precision="mixed", sample_rays=32768, cache={})
    print(f"[autotune] {res['shape']} {res['rays_per_s']:.4g} rays/s", flush=True)
    # slots × workgroups per CU × S in LDS or HBM (mixed, CB on, Mie)
    assert len(res["trials"]) == 16
    assert {t["stage_s"] for t in res["trials"]} == {1, -1}
    assert res["shape"] in [(t["slots"], t["workgroups_per_cu"], t["stage_s"]) for t in res["trials"]]


def test_tuned_batch_shapes_are_applied_and_change_only_rounding():
    """apply_tuned writes the tuned (slots, S placement) into the requests; a
    batch run with them equals the default batch up to FP32 grouping (the same
    rays, events and chunks), and the per-solve shape reaches the device (a
    shape that cannot fit is refused)."""
    from mcrt.tune import apply_tuned
    reqs = _mix()
    ref = []
    with Batch(reqs) as b:
        b.launch()
        b.wait()
        ref = [flat(b.read(i)) for i in range(len(reqs))]
    tuned = [Request(r.cfg, r.g, r.I0, seed=r.seed, precision=r.precision) for r in reqs]
    res = apply_tuned(tuned, sample_chunks=64, cache={})
    print(f"[apply_tuned] {[(r.slots, r.stage_s) for r in tuned]}", flush=True)
    assert all(r.slots in (64, 128, 192, 256) for r in tuned) and res
    forced = [Request(r.cfg, r.g, r.I0, seed=r.seed, precision=r.precision, slots=64, stage_s=-1)
              for r in reqs]
    for rs in (tuned, forced):
        with Batch(rs) as b:
            b.launch()
            b.wait()
            for i, r in enumerate(rs):
                st = {}
                out = flat(b.read(i, stats=st))
                err = float(np.max(np.abs(out - ref[i])))
                tol = (1e-5 if r.precision == "mixed" else 1e-12) * float(np.max(np.abs(ref[i])))
                assert err <= tol, (i, r.slots, r.stage_s, err)


def _expected_chunks(N, tau, omega, bins, cb_bins, rule, min_int=1e-6, maxscat=10000):
    """The default chunking rules as include/mcrt.h states them: batches cut by
    ray count (ceil(N/64) within 512..8192), the server by work (about 4e7 bin
    evaluations per chunk, 64..8192 rays, with the events per ray of a weight
    kept at ω·(1 - p) per event, p = 1/(1 + 0.75·τ^1.15), down to
    minimum_intensity); at least two chunks."""
    import math
    if rule == "latency":
        keep = omega * (1.0 - 1.0 / (1.0 + 0.75 * tau ** 1.15))
        ev = min(maxscat, 1.0 + math.log(1.0 / min_int) / -math.log(keep))
        c = int(max(64.0, min(8192.0, math.ceil(4e7 / (ev * (bins + 3 * cb_bins + 1))))))
    else:
        c = max(512, min(8192, -(-N // 64)))
    if N >= 2 and c >= N:
        c = (N + 1) // 2
    return -(-N // c)


def _chunk_cases():
    from mcrt.batch import BATCH_THETAB_DEG
    from mcrt.config import Config, midpoint_polar_grid
    from mcrt.geometries import SphereGeometry
    from mcrt.scatterers import RayleighScatterer
    thin = W.plane_rayleigh_tau1(4096)          # τ = 1, RT only
    tiny = W.plane_rayleigh_tau1(100)           # fewer rays than one chunk: two chunks
    g = SphereGeometry(R=20.0, scatterer=RayleighScatterer(1.0, 1.0))   # thick, CB on
    ct, st_, w = midpoint_polar_grid(g, 180, 48)
    cfg = Config(g, number_of_rays=1000, maximum_scattering_times=10000, compute_cb=True,
                 Ntheta=180, Nphi=48, costheta=ct, sintheta=st_, inorm=w,
                 thetab=np.deg2rad(BATCH_THETAB_DEG), Nphib=8, minimum_intensity=1e-6, tau0=50.0)
    reqs = [Request(thin[0], thin[1], I0, seed=thin[2], precision="mixed"),
            Request(tiny[0], tiny[1], I0, seed=tiny[2], precision="mixed"),
            Request(cfg, g, I0, seed=3, precision="mixed")]
    args = [(4096, 1.0, 1.0, 180 * 48, 0), (100, 1.0, 1.0, 180 * 48, 0),
            (1000, 20.0, 1.0, 180 * 48, len(BATCH_THETAB_DEG) * 8)]
    return reqs, args


def test_default_chunks_of_batches_and_served_solves():
    from mcrt.serve import Server
    reqs, args = _chunk_cases()
    want_b = [_expected_chunks(*a, rule="throughput") for a in args]
    want_s = [_expected_chunks(*a, rule="latency") for a in args]
    assert want_b == [8, 2, 2] and want_s == [16, 2, 16]
    with Batch(reqs) as b:
        assert b.info()["chunks"] == sum(want_b)
        b.launch()
        b.wait()
        for i, r in enumerate(reqs):
            st = {}
            out = b.read(i, stats=st)
            print(f"[chunks/batch/{i}] chunks={st['chunks']} batches={st['batches']}", flush=True)
            assert st["chunks"] == want_b[i] and st["batches"] == min(128, want_b[i])
            assert st["rays"] == r.cfg.number_of_rays and st["nonfinite"] == 0
            assert abs(sum(out[:5]) - 1.0) < 1e-5
    with Server() as srv:
        tickets = [srv.submit(r) for r in reqs]
        for i, t in enumerate(tickets):
            st = {}
            out = srv.result(t, stats=st)
            print(f"[chunks/server/{i}] chunks={st['chunks']} batches={st['batches']}", flush=True)
            assert st["chunks"] == want_s[i] and st["batches"] == min(128, want_s[i])
            assert st["rays"] == reqs[i].cfg.number_of_rays and st["nonfinite"] == 0
            assert abs(sum(out[:5]) - 1.0) < 1e-5
