"""GPU parity on the configurations the main parity cases leave out: sphere with
Rayleigh scatterers, oblique incidence on a Mie slab, two layers, a
semi-infinite layer, an active τ₀ cut-off, paths cut short by
maximum_scattering_times / minimum_intensity, the reference's default
Gauss-Legendre polar grid, circularly polarised input, slabs whose layers hold
different scatterers (Rayleigh over Mie, Mie over a semi-infinite Rayleigh
layer, Rayleigh over Spline), SplineScatterer spheres and slabs, and Mie
tables too large for the default LDS layout.  Every form runs
the same Philox stream as the oracle, so the criteria are those of
test_gpu_parity (binomial |z| > 3 count and χ²/dof, exact energy telescoping,
event counts within 2%).
"""
import math

import numpy as np
import pytest

import mcrt
import oracle
from mcrt.config import Config
from mcrt.geometries import Layer, PlaneGeometry, SphereGeometry
from mcrt.scatterers import MieScatterer, RayleighScatterer
from statcheck import compare, flat, oracle_flat

pytestmark = pytest.mark.gpu

FORMS = ["fp64_plain", "fp64", "mixed"]


def _mie():
    return MieScatterer.create(0.3, complex(1.45, 0.01), 0.55, 0.1, Ntheta=181, random_points=500)


def _spline():
    # P22 = 0.93·P11, P44 = 0.97·P33 on the Mie tables of _mie()
    return mcrt.workloads.spline_from(_mie_cached())


_MIE = []


def _mie_cached():
    if not _MIE:
        _MIE.append(_mie())
    return _MIE[0]


def _case(name):
    if name == "sphere_rayleigh_cb":
        g = SphereGeometry(R=2.0, scatterer=RayleighScatterer(1.0, 1.0))
        return Config(g, number_of_rays=1024, Ntheta=90, Nphi=16, compute_cb=True), g, 21, (1.0, -1.0, 0.0, 0.0)
    if name == "plane_mie_oblique_cb":
        s = _mie()
        g = PlaneGeometry(math.radians(130.0), math.radians(40.0),
                          [Layer(d=0.8 * s.kl, scatterer=s, tau=0.8, omega=s.omega)])
        return Config(g, number_of_rays=1024, Ntheta=60, Nphi=24, compute_cb=True), g, 22, (1.0, 0.0, 1.0, 0.0)
    if name == "plane_two_layers":
        s = RayleighScatterer(1.0, 1.0)
        g = PlaneGeometry(math.pi, 0.0, [Layer(d=0.5, scatterer=s, tau=0.5, omega=1.0),
                                         Layer(d=1.0, scatterer=s, tau=1.0, omega=0.8)])
        return Config(g, number_of_rays=1024, Ntheta=60, Nphi=24, compute_cb=False), g, 23, (1.0, -1.0, 0.0, 0.0)
    if name == "plane_semi_infinite":
        s = RayleighScatterer(0.8, 1.0)
        g = PlaneGeometry(math.radians(150.0), 0.0, [Layer(d=-1.0, scatterer=s, tau=1.0, omega=0.8)])
        return Config(g, number_of_rays=1024, Ntheta=60, Nphi=24, compute_cb=True), g, 24, (1.0, 1.0, 0.0, 0.0)
    if name == "sphere_mie_tau0_cut":
        s = _mie()
        g = SphereGeometry(R=3.0 * s.kl, scatterer=s, tau=3.0, omega=s.omega)
        return (Config(g, number_of_rays=1024, Ntheta=60, Nphi=16, compute_cb=True, tau0=2.5), g, 25,
                (1.0, 0.0, 0.0, 1.0))
    if name == "plane_rayleigh_over_mie_cb":
        r, m = RayleighScatterer(1.0, 1.0), _mie()
        g = PlaneGeometry(math.pi, 0.0, [Layer(d=0.4, scatterer=r, tau=0.4, omega=1.0),
                                         Layer(d=1.2 * m.kl, scatterer=m, tau=1.2, omega=m.omega)])
        return Config(g, number_of_rays=1024, Ntheta=60, Nphi=24, compute_cb=True), g, 31, (1.0, -1.0, 0.0, 0.0)
    if name == "plane_mie_over_rayleigh_cb":
        r, m = RayleighScatterer(0.9, 1.0), _mie()
        g = PlaneGeometry(math.radians(150.0), math.radians(20.0),
                          [Layer(d=0.5 * m.kl, scatterer=m, tau=0.5, omega=m.omega),
                           Layer(d=-1.0, scatterer=r, tau=1.0, omega=0.9)])
        return Config(g, number_of_rays=1024, Ntheta=60, Nphi=24, compute_cb=True), g, 32, (1.0, 0.0, 1.0, 0.0)
    if name == "plane_rayleigh_over_spline":
        r, sp = RayleighScatterer(1.0, 1.0), _spline()
        g = PlaneGeometry(math.pi, 0.0, [Layer(d=0.3, scatterer=r, tau=0.3, omega=1.0),
                                         Layer(d=sp.kl, scatterer=sp, tau=1.0, omega=sp.omega)])
        return Config(g, number_of_rays=1024, Ntheta=60, Nphi=24, compute_cb=False), g, 33, (1.0, 1.0, 0.0, 0.0)
    if name == "sphere_spline_cb":
        sp = _spline()
        g = SphereGeometry(R=2.0 * sp.kl, scatterer=sp)
        return Config(g, number_of_rays=1024, Ntheta=60, Nphi=16, compute_cb=True), g, 34, (1.0, -1.0, 0.0, 0.0)
    if name == "plane_spline_oblique_cb":
        sp = _spline()
        g = PlaneGeometry(math.radians(140.0), math.radians(70.0),
                          [Layer(d=sp.kl, scatterer=sp, tau=1.0, omega=sp.omega)])
        return Config(g, number_of_rays=1024, Ntheta=60, Nphi=24, compute_cb=True), g, 35, (1.0, 0.0, 0.0, 1.0)
    if name == "plane_three_scatterers_cb":
        # a middle layer whose scatterer is neither table set's: the transport's
        # per-layer tables alone (scatterer(g, idx), cpu.jl:206-234, plane.jl:46-50)
        r, m, sp = RayleighScatterer(0.95, 1.0), _mie(), _spline()
        g = PlaneGeometry(math.radians(145.0), math.radians(30.0),
                          [Layer(d=0.3, scatterer=r, tau=0.3, omega=0.95),
                           Layer(d=0.8 * m.kl, scatterer=m, tau=0.8, omega=m.omega),
                           Layer(d=0.6 * sp.kl, scatterer=sp, tau=0.6, omega=sp.omega)])
        return Config(g, number_of_rays=1024, Ntheta=60, Nphi=24, compute_cb=True), g, 37, (1.0, -1.0, 0.0, 0.0)
    if name == "sphere_mie_2401_cb":
        # beyond the default layout: S from HBM (mixed), 64 slots (fp64)
        s = MieScatterer.create(0.3, complex(1.45, 0.01), 0.55, 0.1, Ntheta=2401, random_points=1000)
        g = SphereGeometry(R=2.0 * s.kl, scatterer=s)
        return Config(g, number_of_rays=512, Ntheta=60, Nphi=16, compute_cb=True), g, 36, (1.0, -1.0, 0.0, 0.0)
    if name == "sphere_spline_2401_cb":
        # a Spline table beyond the default layout: S from HBM (mixed), 64 slots (fp64)
        sp = mcrt.workloads.spline_from(
            MieScatterer.create(0.3, complex(1.45, 0.01), 0.55, 0.1, Ntheta=2401, random_points=1000))
        g = SphereGeometry(R=2.0 * sp.kl, scatterer=sp)
        return Config(g, number_of_rays=512, Ntheta=60, Nphi=16, compute_cb=True), g, 38, (1.0, 0.0, 1.0, 0.0)
    if name == "sphere_short_paths":
        s = _mie()
        g = SphereGeometry(R=4.0 * s.kl, scatterer=s, tau=4.0, omega=s.omega)
        return (Config(g, number_of_rays=1024, Ntheta=60, Nphi=16, compute_cb=True,
                       maximum_scattering_times=3, minimum_intensity=0.05), g, 26, (1.0, -1.0, 0.0, 0.0))
    raise KeyError(name)


CASES = ["sphere_rayleigh_cb", "plane_mie_oblique_cb", "plane_two_layers", "plane_semi_infinite",
         "sphere_mie_tau0_cut", "sphere_short_paths", "plane_rayleigh_over_mie_cb",
         "plane_mie_over_rayleigh_cb", "plane_rayleigh_over_spline", "sphere_spline_cb",
         "plane_spline_oblique_cb", "plane_three_scatterers_cb", "sphere_mie_2401_cb",
         "sphere_spline_2401_cb"]
_ORACLE = {}


def _oracle(name):
    if name not in _ORACLE:
        cfg, g, seed, I0 = _case(name)
        _ORACLE[name] = oracle.Solve(cfg, g).simulate(list(I0), seed, nbatch=32)
    return _ORACLE[name]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("case", CASES)
def test_edge_case_vs_oracle(case, form):
    cfg, g, seed, I0 = _case(case)
    print(f"case={case} form={form} N={cfg.number_of_rays}", flush=True)
    ref = _oracle(case)
    st = {}
    out = mcrt.simulate(cfg, g, list(I0), seed=seed, precision=form, stats=st)
    a = flat(out)
    b = oracle_flat(ref)
    sig = oracle.sigma_of_mean(ref) * np.sqrt(2.0)
    assert abs(sum(out[:5]) - 1.0) < (1e-5 if form == "mixed" else 1e-9)
    assert st["rays"] == cfg.number_of_rays and st["nonfinite"] == 0
    dead_rel = 1e-6 if form == "mixed" else 1e-12
    r = compare(a, b, sig, floor=1e-15, dead_rel=dead_rel, label=f"{case}/{form}")
    assert r["ok"], r
    # same stream: the per-ray histories agree, so every bin agrees far inside σ
    assert r["max_abs_z"] < (1.0 if form == "mixed" else 0.5), r
    ev_ref = ref["counters"][0]
    assert abs(st["events"] - ev_ref) <= 0.02 * ev_ref + 5
    if case == "sphere_short_paths":
        assert st["events"] <= 3 * cfg.number_of_rays


def _big_bins():
    s = MieScatterer.create(0.3, complex(1.45, 0.01), 0.55, 0.1, Ntheta=721, random_points=1000)
    g = SphereGeometry(R=2.0 * s.kl, scatterer=s, tau=2.0, omega=s.omega)
    thetab = np.deg2rad(np.concatenate([np.arange(101) * 0.05, np.arange(6.0, 181.0)]))
    cfg = Config(g, number_of_rays=256, Ntheta=360, Nphi=96, compute_cb=True, thetab=thetab,
                 Nphib=16)
    return cfg, g, 27, (1.0, -1.0, 0.0, 0.0)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_large_bin_grids_vs_oracle(form):
    """34560 RT bins (68 per thread, many register chunks) and 4416 CB bins."""
    cfg, g, seed, I0 = _big_bins()
    if "big" not in _ORACLE:
        _ORACLE["big"] = oracle.Solve(cfg, g).simulate(list(I0), seed, nbatch=32)
    ref = _ORACLE["big"]
    st = {}
    out = mcrt.simulate(cfg, g, list(I0), seed=seed, precision=form, stats=st)
    assert abs(sum(out[:5]) - 1.0) < (1e-5 if form == "mixed" else 1e-9)
    assert st["rays"] == cfg.number_of_rays and st["nonfinite"] == 0
    r = compare(flat(out), oracle_flat(ref), oracle.sigma_of_mean(ref) * np.sqrt(2.0), floor=1e-15,
                dead_rel=1e-6 if form == "mixed" else 1e-12, label=f"big/{form}")
    assert r["ok"], r
    assert r["max_abs_z"] < (1.0 if form == "mixed" else 0.5), r


def test_lds_layout_follows_the_table_size():
    """The cooperative forms keep the scatterer tables in LDS beside the event
    records of the active ray slots: the C2 table (1801 nodes) leaves the full
    slot count with S staged; larger tables first move S to HBM (mixed), then
    take fewer slots; a table that does not fit beside 64 slots is a clear
    MCRT_ERR_UNSUPPORTED, and fp64_plain (tables in device memory) still runs
    it."""
    I0 = [1.0, -1.0, 0.0, 0.0]
    shapes = {}
    for nt in (1801, 2401, 4001, 6001):
        s = MieScatterer.create(0.3, complex(1.45, 0.01), 0.55, 0.1, Ntheta=nt, random_points=1000)
        g = SphereGeometry(R=1.0 * s.kl, scatterer=s)
        cfg = Config(g, number_of_rays=1 << 17, Ntheta=60, Nphi=16, compute_cb=True)
        for form in ("mixed", "fp64"):
            try:
                p = mcrt.Plan(cfg, g, I0, seed=3, precision=form)
            except mcrt._abi.MCRTError as e:
                assert e.code == mcrt._abi.MCRT_ERR_UNSUPPORTED and "LDS" in str(e), e
                shapes[(nt, form)] = None
                continue
            shapes[(nt, form)] = p.shape()
            p.close()
    print(shapes)
    # LDS bytes (160 KiB): F 16·(Nθ+1), S 16·(Nθ+1), csrn 4016, records 336·slots + 128
    # (mixed); F 32·(Nθ+1), csrn 8016, records 640·slots + 128 (fp64)
    assert shapes[(1801, "mixed")]["slots"] == 256 and shapes[(1801, "mixed")]["stage_s"] == 1
    assert shapes[(2401, "mixed")]["slots"] == 256 and shapes[(2401, "mixed")]["stage_s"] == -1
    assert shapes[(4001, "mixed")]["slots"] == 256 and shapes[(4001, "mixed")]["stage_s"] == -1
    assert shapes[(6001, "mixed")]["slots"] == 128
    assert shapes[(1801, "fp64")]["slots"] == 128
    assert shapes[(2401, "fp64")]["slots"] == 64
    assert shapes[(4001, "fp64")] is None  # beyond 64 FP64 slots
    # a Spline table: (F11, F12, F33, F34) and (F22, F44) blocks, 24 (mixed) or 48 (fp64)
    # bytes per node, no csrn in LDS (the general-kind transport reads it from HBM)
    sp = mcrt.workloads.spline_from(
        MieScatterer.create(0.3, complex(1.45, 0.01), 0.55, 0.1, Ntheta=2401, random_points=1000))
    g = SphereGeometry(R=1.0 * sp.kl, scatterer=sp)
    cfg = Config(g, number_of_rays=1 << 17, Ntheta=60, Nphi=16, compute_cb=True)
    sh = {}
    for form in ("mixed", "fp64"):
        p = mcrt.Plan(cfg, g, I0, seed=3, precision=form)
        sh[form] = p.shape()
        p.close()
    print("spline 2401", sh)
    assert sh["mixed"]["slots"] == 256 and sh["mixed"]["stage_s"] == -1
    assert sh["fp64"]["slots"] == 64
    big = MieScatterer.create(0.3, complex(1.45, 0.01), 0.55, 0.1, Ntheta=12001, random_points=1000)
    g = SphereGeometry(R=1.0 * big.kl, scatterer=big)
    cfg = Config(g, number_of_rays=512, Ntheta=60, Nphi=16, compute_cb=True)
    with pytest.raises(mcrt._abi.MCRTError) as e:
        mcrt.simulate(cfg, g, I0, seed=3, precision="mixed")
    assert e.value.code == mcrt._abi.MCRT_ERR_UNSUPPORTED and "LDS" in str(e.value)
    out = mcrt.simulate(cfg, g, I0, seed=3, precision="fp64_plain")
    assert abs(sum(out[:5]) - 1.0) < 1e-9


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_bin_grid_beyond_the_cooperative_limit_is_refused(form):
    """An RT grid of 8192 × 8193 bins (just over 2^26) or a CB grid of that size:
    the cooperative forms refuse it before building anything, where 32-bit bin
    offsets could otherwise wrap (fp64_plain takes any grid that fits in memory)."""
    g = SphereGeometry(R=1.0, scatterer=RayleighScatterer(1.0, 1.0))
    I0 = [1.0, -1.0, 0.0, 0.0]
    for kw in (dict(Ntheta=8192, Nphi=8193, compute_cb=False),
               dict(Ntheta=8, Nphi=8, compute_cb=True, thetab=np.linspace(0.0, 0.1, 8192),
                    Nphib=8193)):
        cfg = Config(g, number_of_rays=64, **kw)
        print(f"form={form} grid={cfg.Ntheta}x{cfg.Nphi} cb={cfg.Nthetab}x{cfg.Nphib}", flush=True)
        with pytest.raises(mcrt._abi.MCRTError) as e:
            mcrt.Plan(cfg, g, I0, seed=1, precision=form)
        assert e.value.code == mcrt._abi.MCRT_ERR_UNSUPPORTED and "2^26" in str(e.value)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["plane_rayleigh_over_mie_cb", "plane_mie_over_rayleigh_cb",
                                  "sphere_spline_cb", "plane_two_layers", "sphere_rayleigh_cb"])
def test_single_ray_histories_match_the_oracle(case, form):
    """Each ray's own history: a plan over one ray (ray_begin = r) makes the
    oracle's trace_ray events (the stream, the rejections and the weights that
    end the ray); in FP64 its energy budget agrees to 1e-9.  In FP32 rounding
    may change a rare history."""
    cfg, g, seed, I0 = _case(case)
    one = cfg.replace(number_of_rays=1)
    S = oracle.Solve(one, g)
    differ = []
    for r in range(16):
        A, _, _, cnt = S.trace_ray(list(I0), seed, r)
        p = mcrt.Plan(one, g, list(I0), seed=seed, precision=form, ray_begin=r)
        try:
            p.launch()
            p.wait()
            st = {}
            out = p.read(stats=st)
        finally:
            p.close()
        if st["events"] != cnt[0] or (form == "fp64" and
                                      np.abs(np.asarray(out[:5]) - A).max() > 1e-9):
            differ.append((r, st["events"], cnt[0]))
    print(f"{case}/{form}: rays whose history differs: {differ}")
    assert len(differ) <= (0 if form == "fp64" else 1), differ
