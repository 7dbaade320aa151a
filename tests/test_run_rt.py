"""run_rt's Mueller assembly (MonteCarloRadiativeTransfer.jl:87-291), pinned on the
CPU through the oracle's transport, and the GPU run_rt against it.

* Sphere, optically thin Rayleigh medium: single scattering dominates, so the
  assembled Mueller matrix must be the Rayleigh matrix of the scattering angle
  Θ (cos Θ = -cos θ at normal incidence): s12/s11 = -(1-μ²)/(1+μ²),
  s22 = s11, s33/s11 = s44/s11 = 2μ/(1+μ²), s34 = 0, s11 ∝ 1+μ².
* Plane: the Q columns are reconstructed exactly from the ±Q runs, and the
  three unpolarised estimates (mean of the ±Q, ±U and ±V runs) agree within
  their Monte Carlo error.
"""
import numpy as np
import pytest

import oracle
from mcrt.api import mueller_plane, mueller_sphere
from mcrt.config import Config
from mcrt.geometries import Layer, PlaneGeometry, SphereGeometry
from mcrt.scatterers import RayleighScatterer
from statcheck import compare


def _run(cfg, g, I, seed, nbatch=16):
    r = oracle.Solve(cfg, g).simulate(list(I), seed, nbatch=nbatch)
    A = r["A"]
    return (A[0], A[1], A[2], A[3], A[4], r["Irt"], r["Icb"]), r


def _thin_sphere(n=20000):
    g = SphereGeometry(R=0.02, scatterer=RayleighScatterer(1.0, 1.0))
    return Config(g, number_of_rays=n, Ntheta=18, Nphi=16, compute_cb=True, Nphib=8), g


def test_sphere_mueller_is_the_rayleigh_matrix_in_the_thin_limit():
    cfg, g = _thin_sphere()
    IS = g.initial_I()
    runs = [(IS[:, i].copy(), _run(cfg, g, IS[:, i], 42 + i)[0]) for i in range(IS.shape[1])]
    M, _, G = mueller_sphere(cfg, runs)
    mu = -cfg.costheta
    s11 = M[0, 0]
    assert np.all(s11 > 0)
    tol = 0.015  # multiple scattering at τ = 0.02 and Monte Carlo error
    assert np.max(np.abs(M[0, 1] / s11 + (1 - mu**2) / (1 + mu**2))) < tol
    assert np.max(np.abs(M[1, 1] / s11 - 1.0)) < tol
    assert np.max(np.abs(M[2, 2] / s11 - 2 * mu / (1 + mu**2))) < tol
    assert np.max(np.abs(M[3, 3] / s11 - 2 * mu / (1 + mu**2))) < tol
    assert np.max(np.abs(M[2, 3] / s11)) < 2e-3
    shape = s11 / (1 + mu**2)
    assert shape.min() / shape.max() > 0.99
    assert abs(G["dt"] + G["abs"] + G["ref"] + G["rem"] - 1.0) < 1e-9


def test_plane_mueller_columns():
    s = RayleighScatterer(1.0, 1.0)
    g = PlaneGeometry(np.pi, 0.0, [Layer(d=0.3, scatterer=s, tau=0.3, omega=1.0)])
    cfg = Config(g, number_of_rays=16000, Ntheta=20, Nphi=8, compute_cb=False)
    IS = g.initial_I()
    out = [_run(cfg, g, IS[:, i], 7 + i, nbatch=32) for i in range(IS.shape[1])]
    runs = [(IS[:, i].copy(), out[i][0]) for i in range(IS.shape[1])]
    M, _, _ = mueller_plane(cfg, runs)
    norm = np.abs(1.0 / (4.0 * cfg.costheta * cfg.inorm))[None, :, None]
    # exact: M·(1, ∓1, 0, 0) is the ∓Q run itself
    for i, q in ((0, -1.0), (1, 1.0)):
        rec = M[:, 0] + q * M[:, 1]
        want = out[i][0][5] * norm
        assert np.allclose(rec, want, rtol=1e-12, atol=1e-15)
    # statistical: the unpolarised response from the ±Q, ±U and ±V pairs
    def pair_mean(i):
        a, b = out[i][1], out[i + 1][1]
        irt = 0.5 * (np.asarray(a["Irt"]) + np.asarray(b["Irt"]))
        sa = oracle.sigma_of_mean(a)[5:5 + irt.size]
        sb = oracle.sigma_of_mean(b)[5:5 + irt.size]
        return irt.transpose(2, 1, 0).ravel(), 0.5 * np.hypot(sa, sb)
    (q, sq), (u, su), (v, sv) = pair_mean(0), pair_mean(2), pair_mean(4)
    for x, sx, name in ((u, su, "U"), (v, sv, "V")):
        r = compare(x, q, np.hypot(sx, sq), floor=1e-15, dead_rel=1e-9, dof=31,
                    label=f"unpolarised Q vs {name}")
        assert r["ok"], r


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_gpu_run_rt_matches_the_oracle_assembly(form):
    import mcrt
    cfg, g = _thin_sphere(8192)
    IS = g.initial_I()
    runs = [(IS[:, i].copy(), _run(cfg, g, IS[:, i], 100 + i)[0]) for i in range(IS.shape[1])]
    M, MCB, _ = mueller_sphere(cfg, runs)
    rt, cb = mcrt.run_rt(cfg, g, use_gpu=True, seed=100, precision=form)
    rtol = 1e-9 if form == "fp64" else 2e-5
    for i in range(4):
        for j in range(4):
            got = np.asarray(rt[f"s{i + 1}{j + 1}"], float)
            scale = float(np.max(np.abs(M[0, 0])))
            err = float(np.max(np.abs(got - M[i, j])))
            print(f"[run_rt/{form}] s{i + 1}{j + 1} max|Δ|={err:.3e}", flush=True)
            assert err <= rtol * scale
            gotb = np.asarray(cb[f"s{i + 1}{j + 1}"], float)
            assert np.allclose(gotb, MCB[i, j], rtol=0, atol=rtol * float(np.max(np.abs(MCB[0, 0]))))
