"""Host logic of the mixed-vs-FP64 precision report (synthetic inputs, no device)."""
import numpy as np

from mcrt import precision
from mcrt import workloads as W


def _fake(cfg, val, sig):
    A = np.full(5, val)
    Irt = np.full((4, cfg.Ntheta, cfg.Nphi), val)
    Icb = np.full((4, cfg.Nthetab, cfg.Nphib), val)
    return (A, Irt, Icb), (np.full(5, sig), np.full_like(Irt, sig), np.full_like(Icb, sig)), {}


def test_group_statistics_scale_with_the_planted_difference():
    cfg, g, seed = W.plane_rayleigh_cb(1000)
    fp64 = _fake(cfg, 1.0, 0.1)
    mixed = _fake(cfg, 1.01, 0.1)
    rep = precision.compare_solves(cfg, fp64, mixed)
    for name in ["A", "Irt_I", "Irt_V", "Icb_I", "Icb_U"]:
        s = rep[name]
        assert np.isclose(s["max_abs_diff"], 0.01)
        assert np.isclose(s["median_diff_over_sigma"], 0.1)
        assert np.isclose(s["max_rel_diff"], 0.01 / 1.0)
    # a uniform 1% offset cancels in the enhancement ratio
    z = rep["cb_enhancement"]
    assert z["max_abs_diff"] < 1e-12
    assert np.isclose(z["theta_b_ref_deg"], 60.0)


def test_enhancement_ratio_and_error():
    Icb = np.ones((4, 3, 2))
    Icb[0, 0] = 2.0   # ζ(0) = 2
    s = np.full_like(Icb, 0.1)
    z, sz = precision.enhancement(Icb, s, None, 2)
    assert np.allclose(z, [2.0, 1.0, 1.0])
    # σ of a 2-azimuth mean: 0.1/√2; ζ(0) error: 2·√((0.0707/2)² + (0.0707/1)²)
    e = 0.1 / np.sqrt(2)
    assert np.isclose(sz[0], 2 * np.sqrt((e / 2) ** 2 + e ** 2))


def test_report_formats_every_group():
    cfg, g, seed = W.plane_rayleigh_cb(1000)
    rep = precision.compare_solves(cfg, _fake(cfg, 1.0, 0.1), _fake(cfg, 1.02, 0.1))
    r = dict(rays=1000, seed=seed, chunk=1000, events=1, events_mixed=1, groups=rep)
    text = precision.format_report(r)
    for name in precision.GROUPS:
        assert f"| {name} |" in text


def test_deterministic_bins_are_reported_apart():
    cfg, g, seed = W.plane_rayleigh_cb(1000)
    fp64 = _fake(cfg, 1.0, 0.1)
    fp64[1][0][0] = 1e-17          # A[0]: the same for every ray
    mixed = _fake(cfg, 1.0, 0.1)
    mixed[0][0][0] = 1.0 + 3e-8
    rep = precision.compare_solves(cfg, fp64, mixed)
    assert rep["A"]["bins_with_sigma"] == 4
    assert np.isclose(rep["A"]["max_rel_diff_deterministic"], 3e-8)
    assert rep["A"]["max_diff_over_sigma"] == 0.0


def test_saved_ranges_merge_by_ray_weight(tmp_path):
    cfg, g, seed = W.plane_rayleigh_cb(1000)
    a = _fake(cfg, 1.0, 0.3)
    b = _fake(cfg, 3.0, 0.4)
    pa, pb = tmp_path / "a.npz", tmp_path / "b.npz"
    precision.save_solve(pa, cfg, a, ray_begin=0, rays=100, form="fp64", seed=seed)
    precision.save_solve(pb, cfg, b, ray_begin=100, rays=300, form="fp64", seed=seed)
    (vals, sig, st), N, ranges = precision.load_merged([pb, pa])
    assert N == 400 and ranges == [(0, 100), (100, 300)]
    assert np.allclose(vals[1], 0.25 * 1.0 + 0.75 * 3.0)
    assert np.allclose(sig[1], np.hypot(0.25 * 0.3, 0.75 * 0.4))
    precision.save_solve(pb, cfg, b, ray_begin=150, rays=300, form="fp64", seed=seed)
    import pytest
    with pytest.raises(ValueError):
        precision.load_merged([pa, pb])
