"""Scatterers: RayleighScatterer, MieScatterer and SplineScatterer.

Mirrors src/scatterers/index.jl (mean_free_path, single_scattering_albedo),
src/scatterers/rayleigh.jl:3-6, src/scatterers/mie.jl:3-44 and
src/scatterers/spline.jl:3-24.  The Mie tables come from the library's FP64
builders (``mcrt_bhmie``, ``mcrt_calculate_csrn``), the same arithmetic the
device tables are staged from.  SplineScatterer's lookups read ``self.NSPL``
in the reference, a name defined nowhere (every call throws); the library
reads it as Nθ - 1, the table's last node, and keeps the rest as written.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi


class AbstractScatterer:
    """Base of the scatterer types (MonteCarloRadiativeTransfer.jl:16)."""

    omega: float
    kl: float


def mean_free_path(s: AbstractScatterer) -> float:
    """mean_free_path(s) = s.kl (scatterers/index.jl:1)."""
    return s.kl


def single_scattering_albedo(s: AbstractScatterer) -> float:
    """single_scattering_albedo(s) = s.ω (scatterers/index.jl:3)."""
    return s.omega


@dataclass(frozen=True)
class RayleighScatterer(AbstractScatterer):
    """RayleighScatterer(ω, kl) — src/scatterers/rayleigh.jl:3-6."""

    omega: float
    kl: float


@dataclass(frozen=True)
class BHMieResult:
    """Output of bhmie (src/scatterers/mie.jl:214-215)."""

    q_ext: float
    q_sca: float
    q_abs: float
    q_back: float
    asymm: float
    S1: np.ndarray  # complex128 [Nθ]
    S2: np.ndarray
    F11: np.ndarray
    F12: np.ndarray
    F33: np.ndarray
    F34: np.ndarray

    @property
    def S(self):
        return (self.S1, self.S2)

    @property
    def F(self):
        return (self.F11, self.F12, self.F33, self.F34)


def bhmie(x: float, m: complex, Ntheta: int = 181, nextra: int = 15, custom_nstop: int = 0) -> BHMieResult:
    """bhmie(x, m; Nθ, nextra, custom_nstop) — src/scatterers/mie.jl:129-216 (FP64)."""
    lib = _abi.load_library()
    m = complex(m)
    N = int(Ntheta)
    q = np.zeros(5)
    S1 = np.zeros(2 * N)
    S2 = np.zeros(2 * N)
    F = [np.zeros(N) for _ in range(4)]
    rc = lib.mcrt_bhmie(float(x), m.real, m.imag, N, int(nextra), int(custom_nstop), _abi.dptr(q),
                        _abi.dptr(S1), _abi.dptr(S2), *[_abi.dptr(f) for f in F])
    _abi.check(rc)
    return BHMieResult(q[0], q[1], q[2], q[3], q[4], S1[0::2] + 1j * S1[1::2],
                       S2[0::2] + 1j * S2[1::2], *F)


def calculate_csrn(F11: np.ndarray, random_points: int) -> np.ndarray:
    """calculate_csrn(linear_interpolation(range(0, π, Nθ), F11), n) — src/utils.jl:263-287."""
    lib = _abi.load_library()
    F11 = _abi.f64(F11)
    out = np.zeros(int(random_points) + 1)
    _abi.check(lib.mcrt_calculate_csrn(_abi.dptr(F11), len(F11), int(random_points), _abi.dptr(out)))
    return out


@dataclass(frozen=True, eq=False)
class MieScatterer(AbstractScatterer):
    """MieScatterer — src/scatterers/mie.jl:3-15 (struct) and :36-44 (constructor).

    Build it with ``MieScatterer.create(r, m, λ, ν; Ntheta, random_points)``
    (the reference's ``MieScatterer(r, m, λ, ν; Nθ, random_points)``), or pass
    the fields directly.
    """

    omega: float
    kl: float
    csrn: np.ndarray
    F11: np.ndarray
    F12: np.ndarray
    F33: np.ndarray
    F34: np.ndarray
    S11: np.ndarray  # complex128
    S22: np.ndarray
    Ntheta: int
    random_points: int
    params: dict = field(default_factory=dict)

    @classmethod
    def create(cls, r: float, m: complex, lam: float, nu: float, Ntheta: int = 181,
               random_points: int = 1000) -> "MieScatterer":
        x = 2 * math.pi * r / lam
        res = bhmie(x, m, Ntheta=Ntheta)
        omega = res.q_sca / res.q_ext
        kl = ((4 / 3) * r) / (nu * res.q_ext)
        csrn = calculate_csrn(res.F11, random_points)
        return cls(omega, kl, csrn, res.F11, res.F12, res.F33, res.F34, res.S1, res.S2, int(Ntheta),
                   int(random_points), {"r": r, "m": complex(m), "lam": lam, "nu": nu, "x": x,
                                        "q_ext": res.q_ext, "q_sca": res.q_sca,
                                        "asymm": res.asymm})


@dataclass(frozen=True, eq=False)
class SplineScatterer(AbstractScatterer):
    """SplineScatterer — src/scatterers/spline.jl:3-15 (struct) and :19-24
    (constructor): user tables F11, F12, F22, F33, F34, F44, S11, S22 of one
    length Nθ on θ = range(0, π, Nθ), ω and kl given, csrn built from F11.

    Build it with ``SplineScatterer.create(ω, kl, F11, F12, F22, F33, F34,
    F44, S11, S22; random_points)`` (the reference's positional constructor),
    or pass the fields directly.  The lookups (spline.jl:26-76) take
    NSPL = Nθ - 1; the phase matrix keeps the reference's weights
    (msn+1-sn, sn-msn), one node step on from a linear interpolation.
    """

    omega: float
    kl: float
    csrn: np.ndarray
    F11: np.ndarray
    F12: np.ndarray
    F22: np.ndarray
    F33: np.ndarray
    F34: np.ndarray
    F44: np.ndarray
    S11: np.ndarray  # complex128
    S22: np.ndarray

    @classmethod
    def create(cls, omega: float, kl: float, F11, F12, F22, F33, F34, F44, S11, S22,
               random_points: int = 1000) -> "SplineScatterer":
        F = [_abi.f64(f) for f in (F11, F12, F22, F33, F34, F44)]
        S = [np.asarray(z, dtype=np.complex128) for z in (S11, S22)]
        if any(len(a) != len(F[0]) for a in F + S):
            raise ValueError("the Spline tables must have one length")
        csrn = calculate_csrn(F[0], random_points)
        return cls(float(omega), float(kl), csrn, *F, *S)

    @property
    def Ntheta(self) -> int:
        return len(self.F11)

    @property
    def random_points(self) -> int:
        return len(self.csrn) - 1


def to_struct(s: AbstractScatterer, keep: list) -> _abi.mcrt_scatterer:
    """Fill an mcrt_scatterer; arrays referenced by it are appended to keep."""
    st = _abi.mcrt_scatterer()
    st.omega = float(s.omega)
    st.kl = float(s.kl)
    if isinstance(s, RayleighScatterer):
        st.kind = _abi.SCATTERER_RAYLEIGH
        return st
    if isinstance(s, MieScatterer):
        st.kind = _abi.SCATTERER_MIE
        names = ("csrn", "F11", "F12", "F33", "F34")
    elif isinstance(s, SplineScatterer):
        st.kind = _abi.SCATTERER_SPLINE
        names = ("csrn", "F11", "F12", "F22", "F33", "F34", "F44")
    else:
        raise TypeError(f"unsupported scatterer type {type(s).__name__}")
    st.n_theta = int(s.Ntheta)
    st.random_points = int(s.random_points)
    arrs = {k: _abi.f64(getattr(s, k)) for k in names}
    for k in ("S11", "S22"):
        z = np.asarray(getattr(s, k), dtype=np.complex128)
        arrs[k] = np.ascontiguousarray(np.stack([z.real, z.imag], axis=1).reshape(-1))
    if len(arrs["csrn"]) != st.random_points + 1:
        raise ValueError("csrn must hold random_points + 1 values")
    for k in names[1:] + ("S11", "S22"):
        if len(arrs[k]) != st.n_theta * (2 if k[0] == "S" else 1):
            raise ValueError(f"{k} must hold Nθ values")
    for k, a in arrs.items():
        keep.append(a)
        setattr(st, k, _abi.dptr(a))
    return st
