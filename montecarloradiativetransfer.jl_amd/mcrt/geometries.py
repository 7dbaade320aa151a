"""Geometries: Layer, PlaneGeometry and SphereGeometry.

Mirrors src/geometries/plane.jl and src/geometries/sphere.jl: the structs,
``validate``, ``initial_I`` and ``initial_weights``.  The per-ray queries
(distance_to_boundary, pass_boundary, …) live in the device code.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _abi
from .scatterers import AbstractScatterer, mean_free_path, single_scattering_albedo, to_struct


class AbstractGeometry:
    """Base of the geometry types (MonteCarloRadiativeTransfer.jl:17)."""


class ArgumentError(ValueError):
    """The reference's ArgumentError from validate."""


_UNSET = object()


@dataclass
class Layer:
    """Layer(d, n=1, scatterer, τ=d/mean_free_path, ω=albedo) — plane.jl:6-21.

    d < 0 marks a semi-infinite layer (Base.isinf(::Layer), plane.jl:23)."""

    d: float
    scatterer: AbstractScatterer
    n: float = 1.0
    tau: float = None  # type: ignore[assignment]
    omega: float = None  # type: ignore[assignment]

    def __post_init__(self):
        if self.tau is None:
            self.tau = self.d / mean_free_path(self.scatterer)
        if self.omega is None:
            self.omega = single_scattering_albedo(self.scatterer)

    def isinf(self) -> bool:
        return self.d < 0


@dataclass
class PlaneGeometry(AbstractGeometry):
    """PlaneGeometry(θ, ϕ, layers) — plane.jl:28-38."""

    theta: float
    phi: float
    layers: Sequence[Layer]

    def validate(self) -> None:  # plane.jl:52-62
        if not (math.pi / 2 < self.theta <= math.pi):
            raise ArgumentError("θ must be in (π/2, π]")
        if not (0 <= self.phi < 2 * math.pi):
            raise ArgumentError("ϕ must be in [0, 2π)")
        if len(self.layers) == 0:
            raise ArgumentError("There needs to be at least one layer")
        for i, l in enumerate(self.layers):
            if i != len(self.layers) - 1 and l.isinf():
                raise ArgumentError("Only the last layer can be infinite")

    def incidence_angle(self):
        return (self.theta, self.phi)

    @staticmethod
    def initial_I() -> np.ndarray:  # plane.jl:70-75 (columns are incident Stokes vectors)
        return np.array([[1.0, 1.0, 1.0, 1.0, 1.0, 1.0],
                         [-1.0, 1.0, 0.0, 0.0, 0.0, 0.0],
                         [0.0, 0.0, -1.0, 1.0, 0.0, 0.0],
                         [0.0, 0.0, 0.0, 0.0, -1.0, 1.0]])


@dataclass
class SphereGeometry(AbstractGeometry):
    """SphereGeometry(R, n=1, scatterer, τ=R/mean_free_path, ω=albedo) — sphere.jl:6-21."""

    R: float
    scatterer: AbstractScatterer
    n: float = 1.0
    tau: float = None  # type: ignore[assignment]
    omega: float = None  # type: ignore[assignment]

    def __post_init__(self):
        if self.tau is None:
            self.tau = self.R / mean_free_path(self.scatterer)
        if self.omega is None:
            self.omega = single_scattering_albedo(self.scatterer)

    def validate(self) -> None:  # sphere.jl:35-37
        if not self.R > 0:
            raise ArgumentError("Radius must be positive")

    def incidence_angle(self):  # sphere.jl:41
        return (math.pi, 0.0)

    @staticmethod
    def initial_I() -> np.ndarray:  # sphere.jl:49-52
        return np.array([[1.0, 1.0], [-1.0, 0.0], [0.0, 0.0], [0.0, -1.0]])


def gausslegendre(n: int):
    """FastGaussQuadrature.gausslegendre(n) via the library's FP64 builder."""
    lib = _abi.load_library()
    x = np.zeros(int(n))
    w = np.zeros(int(n))
    _abi.check(lib.mcrt_gauss_legendre(int(n), _abi.dptr(x), _abi.dptr(w)))
    return x, w


def initial_weights(g: AbstractGeometry, Ntheta: int, Nphi: int):
    """initial_weights(g, Nθ, Nϕ) → (cosθ, sinθ, inorm) — plane.jl:131-142, sphere.jl:85-90."""
    lib = _abi.load_library()
    kind = _abi.GEOMETRY_PLANE if isinstance(g, PlaneGeometry) else _abi.GEOMETRY_SPHERE
    ct, st, w = (np.zeros(int(Ntheta)) for _ in range(3))
    _abi.check(lib.mcrt_initial_weights(kind, int(Ntheta), int(Nphi), _abi.dptr(ct), _abi.dptr(st),
                                        _abi.dptr(w)))
    return ct, st, w


def to_struct(g: AbstractGeometry, keep: list) -> _abi.mcrt_geometry:
    """Fill an mcrt_geometry (scatterers are shared between layers by identity)."""
    from .scatterers import to_struct as sca_struct

    gs = _abi.mcrt_geometry()
    cache: dict[int, object] = {}

    def sptr(s):
        if id(s) not in cache:
            st = sca_struct(s, keep)
            keep.append(st)
            cache[id(s)] = st
        import ctypes as C
        return C.pointer(cache[id(s)])

    if isinstance(g, PlaneGeometry):
        gs.kind = _abi.GEOMETRY_PLANE
        gs.theta = float(g.theta)
        gs.phi = float(g.phi)
        arr = (_abi.mcrt_layer * max(1, len(g.layers)))()
        for i, l in enumerate(g.layers):
            arr[i].d = float(l.d)
            arr[i].n = float(l.n)
            arr[i].tau = float(l.tau)
            arr[i].omega = float(l.omega)
            arr[i].scatterer = sptr(l.scatterer)
        keep.append(arr)
        gs.n_layers = len(g.layers)
        gs.layers = arr
    elif isinstance(g, SphereGeometry):
        gs.kind = _abi.GEOMETRY_SPHERE
        gs.R = float(g.R)
        gs.n = float(g.n)
        gs.tau = float(g.tau)
        gs.omega = float(g.omega)
        gs.scatterer = sptr(g.scatterer)
    else:
        raise TypeError(f"unsupported geometry {type(g).__name__}")
    return gs
