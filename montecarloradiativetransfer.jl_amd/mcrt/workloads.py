"""Named workloads of the benchmark and the parity tests.

  plane_rayleigh_tau1   plane slab τ=1, ω=1, Rayleigh, normal incidence, RT only,
                        ≤1000 orders, 180 polar bins of 1° × 48 azimuths, seed 42
  sphere_mie_cb         sphere of optical radius 5, Mie spheres r=0.5 µm at
                        λ=0.55 µm (x≈5.712), m=1.5+0.001i, ν=0.1, F/S tabulated on
                        1801 angles (0.1°), random_points=1000; RT 180 × 1° polar
                        bins × 48 azimuths; CB θb = 0:0.05°:5° (101) × 8; seed 7
  plane_rayleigh_cb     plane_rayleigh_tau1 with CB on (default θb grid), 1e9 rays:
                        the mixed-vs-FP64 precision case
  sphere_mie_rt         sphere_mie_cb without CB
  sphere_spline_cb      sphere_mie_cb with a SplineScatterer on the same tables,
                        P22 = 0.93·P11 and P44 = 0.97·P33 (a depolarising
                        scatterer the Mie form cannot express)
  plane_rayleigh_over_mie  two layers: Rayleigh τ = 0.5 over the C2 Mie
                        scatterer τ = 1.5, normal incidence, CB on
  plane_mie_cb          plane slab τ = 1 of the C2 Mie scatterer, normal incidence,
                        sphere_mie_cb's RT and CB bins, 1e6 rays, seed 13
The ray counts of the full workloads are 1e6, 1e7 and 1e9; tests use fewer rays
of the same configuration.
"""
from __future__ import annotations

import math

import numpy as np

from .config import Config, midpoint_polar_grid
from .geometries import Layer, PlaneGeometry, SphereGeometry
from .scatterers import MieScatterer, RayleighScatterer, SplineScatterer

_MIE_CACHE: dict = {}


def c2_mie(Ntheta: int = 1801) -> MieScatterer:
    key = ("c2", Ntheta)
    if key not in _MIE_CACHE:
        _MIE_CACHE[key] = MieScatterer.create(0.5, 1.5 + 0.001j, 0.55, 0.1, Ntheta=Ntheta,
                                              random_points=1000)
    return _MIE_CACHE[key]


def spline_from(s: MieScatterer, f22: float = 0.93, f44: float = 0.97) -> SplineScatterer:
    """A SplineScatterer on a Mie scatterer's tables with P22 = f22·P11 and
    P44 = f44·P33 (ω, kl and the amplitude tables as the Mie scatterer's)."""
    key = ("spline", id(s), f22, f44)
    if key not in _MIE_CACHE:
        _MIE_CACHE[key] = SplineScatterer.create(s.omega, s.kl, s.F11, s.F12, f22 * s.F11, s.F33,
                                                 s.F34, f44 * s.F33, s.S11, s.S22,
                                                 random_points=s.random_points)
    return _MIE_CACHE[key]


def plane_rayleigh_tau1(n_rays: int = 10**6, *, compute_cb: bool = False, gl_grid: bool = False,
                        max_orders: int = 1000, tau: float = 1.0, omega: float = 1.0,
                        Nphi: int = 48, Ntheta: int = 180):
    """Plane-parallel slab, Rayleigh, normal incidence (θ = π)."""
    sca = RayleighScatterer(omega, 1.0)
    g = PlaneGeometry(math.pi, 0.0, [Layer(d=tau * sca.kl, scatterer=sca, tau=tau, omega=omega)])
    kw = dict(number_of_rays=n_rays, maximum_scattering_times=max_orders, compute_cb=compute_cb,
              Ntheta=Ntheta, Nphi=Nphi, minimum_intensity=1e-6, tau0=50.0)
    if not gl_grid:
        ct, st, w = midpoint_polar_grid(g, Ntheta, Nphi)
        kw.update(costheta=ct, sintheta=st, inorm=w)
    cfg = Config(g, **kw)
    return cfg, g, 42


def sphere_mie_cb(n_rays: int = 10**7, *, Ntheta_table: int = 1801, Nphib: int = 8,
                  thetab_deg=None, compute_cb: bool = True, max_orders: int = 10000,
                  tau: float = 5.0, Nphi: int = 48, Ntheta: int = 180):
    """Sphere of Mie scatterers with CB near exact backscattering."""
    sca = c2_mie(Ntheta_table)
    g = SphereGeometry(R=tau * sca.kl, scatterer=sca)
    if thetab_deg is None:
        thetab_deg = np.arange(101) * 0.05
    ct, st, w = midpoint_polar_grid(g, Ntheta, Nphi)
    cfg = Config(g, number_of_rays=n_rays, maximum_scattering_times=max_orders,
                 compute_cb=compute_cb, Ntheta=Ntheta, Nphi=Nphi, costheta=ct, sintheta=st,
                 inorm=w, thetab=np.deg2rad(np.asarray(thetab_deg, dtype=float)), Nphib=Nphib,
                 minimum_intensity=1e-6, tau0=50.0)
    return cfg, g, 7


def plane_rayleigh_cb(n_rays: int = 10**9):
    """The plane slab of plane_rayleigh_tau1 with CB on (the reference's default
    θb grid of 51 angles × 8 azimuths, kl = 1): the mixed-vs-FP64 precision case."""
    return plane_rayleigh_tau1(n_rays, compute_cb=True)


def sphere_mie_rt(n_rays: int = 10**7):
    """sphere_mie_cb without CB (the peel-off alone)."""
    return sphere_mie_cb(n_rays, compute_cb=False)


def sphere_spline_cb(n_rays: int = 10**7, *, tau: float = 5.0):
    """sphere_mie_cb with a SplineScatterer on the C2 tables (spline_from)."""
    cfg, g, seed = sphere_mie_cb(n_rays, tau=tau)
    sp = spline_from(c2_mie(1801))
    g2 = SphereGeometry(R=tau * sp.kl, scatterer=sp)
    return cfg.replace(geometry=g2), g2, seed


def plane_rayleigh_over_mie(n_rays: int = 10**6, *, tau_top: float = 0.5, tau_mie: float = 1.5,
                            compute_cb: bool = True, Ntheta: int = 180, Nphi: int = 48):
    """Two layers with different scatterers: Rayleigh (kl = 1) over the C2 Mie
    scatterer, normal incidence.  The peel-off's reflected range uses the
    Rayleigh layer, the transmitted range the Mie layer, CB layer 1's
    scatterer (cpu.jl:50)."""
    ray = RayleighScatterer(1.0, 1.0)
    mie = c2_mie(1801)
    g = PlaneGeometry(math.pi, 0.0, [Layer(d=tau_top * ray.kl, scatterer=ray, tau=tau_top, omega=1.0),
                                     Layer(d=tau_mie * mie.kl, scatterer=mie, tau=tau_mie,
                                           omega=mie.omega)])
    ct, st, w = midpoint_polar_grid(g, Ntheta, Nphi)
    cfg = Config(g, number_of_rays=n_rays, maximum_scattering_times=10000, compute_cb=compute_cb,
                 Ntheta=Ntheta, Nphi=Nphi, costheta=ct, sintheta=st, inorm=w,
                 minimum_intensity=1e-6, tau0=50.0)
    return cfg, g, 11


def plane_mie_cb(n_rays: int = 10**6, *, tau: float = 1.0, compute_cb: bool = True):
    """A plane slab of the C2 Mie scatterer at normal incidence with CB on the
    sphere workload's bins (RT 180 × 48 midpoint bins, θb 0:0.05°:5° × 8)."""
    mie = c2_mie(1801)
    g = PlaneGeometry(math.pi, 0.0, [Layer(d=tau * mie.kl, scatterer=mie, tau=tau, omega=mie.omega)])
    ct, st, w = midpoint_polar_grid(g, 180, 48)
    cfg = Config(g, number_of_rays=n_rays, maximum_scattering_times=10000, compute_cb=compute_cb,
                 Ntheta=180, Nphi=48, costheta=ct, sintheta=st, inorm=w,
                 thetab=np.deg2rad(np.arange(101) * 0.05), Nphib=8, minimum_intensity=1e-6,
                 tau0=50.0)
    return cfg, g, 13


WORKLOADS = {"plane_rayleigh_tau1": plane_rayleigh_tau1, "sphere_mie_cb": sphere_mie_cb,
             "plane_rayleigh_cb": plane_rayleigh_cb, "sphere_mie_rt": sphere_mie_rt,
             "sphere_spline_cb": sphere_spline_cb, "plane_rayleigh_over_mie": plane_rayleigh_over_mie,
             "plane_mie_cb": plane_mie_cb}
