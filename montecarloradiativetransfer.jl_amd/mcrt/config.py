"""Config — src/config.jl:55-124.

Field names are the reference's with ASCII spellings (τ₀ → tau0, Nθ → Ntheta,
ϕ → phi, θb → thetab, …); the Julia spellings are accepted as keyword aliases,
e.g. ``Config(geometry=g, **{"Nθ": 180, "τ₀": 50.0})``.  As in the reference,
the polar bins double as quadrature nodes and weights (config.jl:92-99).
"""
from __future__ import annotations

import math

import numpy as np

from . import _abi
from .geometries import AbstractGeometry, initial_weights

# DEFAULT_THEB, config.jl:3-53 (degrees; converted to radians)
DEFAULT_THEB_DEG = np.array([
    0.00, 0.02, 0.04, 0.06, 0.08, 0.10, 0.15, 0.20, 0.25, 0.30, 0.35, 0.40, 0.45, 0.50, 0.60, 0.70,
    0.80, 0.90, 1.00, 1.50, 2.00, 2.50, 3.00, 4.00, 5.00, 6.00, 7.00, 8.00, 9.00, 12.0, 15.0, 18.0,
    21.0, 24.0, 27.0, 30.0, 40.0, 50.0, 60.0, 70.0, 80.0, 90.0, 100.0, 110.0, 120.0, 130.0, 140.0,
    150.0, 160.0, 170.0, 180.0])
DEFAULT_THEB = math.pi / 180.0 * DEFAULT_THEB_DEG

ALIASES = {
    "τ₀": "tau0", "τ0": "tau0", "Nϕ": "Nphi", "Nφ": "Nphi", "ϕ": "phi", "φ": "phi",
    "cosϕ": "cosphi", "cosφ": "cosphi", "sinϕ": "sinphi", "sinφ": "sinphi", "Nθ": "Ntheta",
    "cosθ": "costheta", "sinθ": "sintheta", "θb": "thetab", "Nθb": "Nthetab", "cosθb": "costhetab",
    "sinθb": "sinthetab", "Nϕb": "Nphib", "Nφb": "Nphib", "ϕb": "phib", "φb": "phib",
    "cosϕb": "cosphib", "cosφb": "cosphib", "sinϕb": "sinphib", "sinφb": "sinphib",
}

_FIELDS = ["geometry", "number_of_rays", "minimum_intensity", "maximum_scattering_times",
           "maximum_threads", "compute_cb", "tau0", "Nphi", "phi", "cosphi", "sinphi", "Ntheta",
           "costheta", "sintheta", "inorm", "thetab", "Nthetab", "costhetab", "sinthetab", "Nphib",
           "phib", "cosphib", "sinphib"]


class Config:
    """All run parameters of a solve (config.jl:55-124), with the same defaults."""

    def __init__(self, geometry: AbstractGeometry, **kw):
        kw = {ALIASES.get(k, k): v for k, v in kw.items()}
        unknown = set(kw) - set(_FIELDS)
        if unknown:
            raise TypeError(f"Config: unknown field(s) {sorted(unknown)}")
        self.geometry = geometry
        self.number_of_rays = int(kw.get("number_of_rays", 1000))
        self.minimum_intensity = float(kw.get("minimum_intensity", 1e-6))
        self.maximum_scattering_times = int(kw.get("maximum_scattering_times", 10000))
        self.maximum_threads = int(kw.get("maximum_threads", 8192))
        self.compute_cb = bool(kw.get("compute_cb", True))
        self.tau0 = float(kw.get("tau0", 50.0))
        self.Nphi = int(kw.get("Nphi", 48))
        self.phi = _abi.f64(kw["phi"]) if "phi" in kw else np.linspace(
            0.0, 2 * math.pi * (self.Nphi - 1) / self.Nphi, self.Nphi)
        self.cosphi = _abi.f64(kw["cosphi"]) if "cosphi" in kw else np.cos(self.phi)
        self.sinphi = _abi.f64(kw["sinphi"]) if "sinphi" in kw else np.sin(self.phi)
        self.Ntheta = int(kw.get("Ntheta", 64))
        if any(k not in kw for k in ("costheta", "sintheta", "inorm")):
            ct, st, w = initial_weights(geometry, self.Ntheta, self.Nphi)
        self.costheta = _abi.f64(kw["costheta"]) if "costheta" in kw else ct
        self.sintheta = _abi.f64(kw["sintheta"]) if "sintheta" in kw else st
        self.inorm = _abi.f64(kw["inorm"]) if "inorm" in kw else w
        self.thetab = _abi.f64(kw.get("thetab", DEFAULT_THEB))
        self.Nthetab = int(kw.get("Nthetab", len(self.thetab)))
        self.costhetab = _abi.f64(kw["costhetab"]) if "costhetab" in kw else np.cos(self.thetab)
        self.sinthetab = _abi.f64(kw["sinthetab"]) if "sinthetab" in kw else np.sin(self.thetab)
        self.Nphib = int(kw.get("Nphib", 8))
        self.phib = _abi.f64(kw["phib"]) if "phib" in kw else np.linspace(
            0.0, 2 * math.pi * (self.Nphib - 1) / self.Nphib, self.Nphib)
        self.cosphib = _abi.f64(kw["cosphib"]) if "cosphib" in kw else np.cos(self.phib)
        self.sinphib = _abi.f64(kw["sinphib"]) if "sinphib" in kw else np.sin(self.phib)
        for name, n in (("phi", self.Nphi), ("cosphi", self.Nphi), ("sinphi", self.Nphi),
                        ("costheta", self.Ntheta), ("sintheta", self.Ntheta), ("inorm", self.Ntheta),
                        ("costhetab", self.Nthetab), ("sinthetab", self.Nthetab),
                        ("cosphib", self.Nphib), ("sinphib", self.Nphib)):
            if len(getattr(self, name)) != n:
                raise ValueError(f"Config: {name} must have {n} entries")

    def replace(self, **kw) -> "Config":
        """A copy with some fields replaced (the reference's Config(; …, field=v))."""
        d = {k: getattr(self, k) for k in _FIELDS if k != "geometry"}
        kw = {ALIASES.get(k, k): v for k, v in kw.items()}
        geometry = kw.pop("geometry", self.geometry)
        # derived vectors follow their counts unless given
        if "Ntheta" in kw:
            for k in ("costheta", "sintheta", "inorm"):
                d.pop(k)
        if "Nphi" in kw:
            for k in ("phi", "cosphi", "sinphi"):
                d.pop(k)
            if "Ntheta" not in kw:
                d.pop("inorm", None); d.pop("costheta", None); d.pop("sintheta", None)
        if "phi" in kw:
            d.pop("cosphi"); d.pop("sinphi")
        if "thetab" in kw:
            for k in ("Nthetab", "costhetab", "sinthetab"):
                d.pop(k)
        if "Nphib" in kw:
            for k in ("phib", "cosphib", "sinphib"):
                d.pop(k)
        if "phib" in kw:
            for k in ("Nphib", "cosphib", "sinphib"):
                d.pop(k, None)
            kw.setdefault("Nphib", len(kw["phib"]))
        d.update(kw)
        return Config(geometry, **d)

    def to_struct(self, keep: list) -> _abi.mcrt_config:
        c = _abi.mcrt_config()
        c.number_of_rays = self.number_of_rays
        c.minimum_intensity = self.minimum_intensity
        c.maximum_scattering_times = self.maximum_scattering_times
        c.maximum_threads = self.maximum_threads
        c.compute_cb = 1 if self.compute_cb else 0
        c.tau0 = self.tau0
        c.n_phi = self.Nphi
        c.n_theta = self.Ntheta
        c.n_theta_b = self.Nthetab
        c.n_phi_b = self.Nphib
        for fld, arr in (("phi", self.phi), ("cos_phi", self.cosphi), ("sin_phi", self.sinphi),
                         ("cos_theta", self.costheta), ("sin_theta", self.sintheta),
                         ("inorm", self.inorm), ("theta_b", self.thetab),
                         ("cos_theta_b", self.costhetab), ("sin_theta_b", self.sinthetab),
                         ("phi_b", self.phib), ("cos_phi_b", self.cosphib),
                         ("sin_phi_b", self.sinphib)):
            a = _abi.f64(arr)
            keep.append(a)
            setattr(c, fld, _abi.dptr(a))
        return c


def midpoint_polar_grid(geometry: AbstractGeometry, Ntheta: int, Nphi: int):
    """A uniform polar grid of Ntheta bins of width π/Ntheta as (cosθ, sinθ, inorm).

    Nodes are bin midpoints ordered by increasing cosθ (the transmission half
    first for a plane, as lower_theta_range = 1:Nθ÷2 expects), and
    inorm = |Δcosθ|/(2Nϕ) (Δcosθ over the bin: ΔΩ/4π per azimuth bin), which
    is the normalisation of the Gauss–Legendre weights in plane.jl:136-139 and
    sphere.jl:88, so Σ inorm·Nϕ = 1.
    """
    edges = np.linspace(math.pi, 0.0, Ntheta + 1)  # θ from π down to 0
    mid = 0.5 * (edges[:-1] + edges[1:])
    ct = np.cos(mid)
    st = np.sqrt(1.0 - ct * ct)  # as initial_weights forms sinθ
    w = np.abs(np.cos(edges[1:]) - np.cos(edges[:-1])) / (2 * Nphi)
    return ct, st, w
