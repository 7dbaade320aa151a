"""Frozen algorithmic operation counts of one bin evaluation, as written in the
reference (the unit of work of the roofline figures).

Convention: every add, sub, mul, div and sqrt counts 1; each transcendental
(exp, acos, sin, cos, cbrt, log) counts 1; negations, compares and integer
index arithmetic count 0; dense StaticArrays products count their zero terms
(as written).  Counts are taken from the Julia source:

  get_basis            utils.jl:52-62    k₃², 1-, √, 1/, 4 mul                =   8
  norm_to_ray          utils.jl:41-45    10 mul + 4 add (row 1-3: 7, 3, 7)     =  17
  ray_to_norm          utils.jl:30-34                                           =  17
  spherical_from_cart  utils.jl:9-23     s² 3, sinθ 3, 1/√s² 2, cosϕ sinϕ 2     =  10
  update_I frame part  utils.jl:92-105   get_basis + norm_to_ray + sph + Ehk 17
                                         + Evk (2 + 17) + cosψ/sinψ 6           =  77
  rotate + angles      utils.jl:108-109  (3 + 2 + 6) × 2                        =  22
  dense 4×4·4          utils.jl:108      16 mul + 12 add                        =  28
  phase matrix         rayleigh.jl:12-17 μ², 1±μ² (2), ×0.75 (2), 1.5μ         =   6
                       mie.jl:53-71      acos, /π, ×(N-1), w 3, 8 mul, 4 add    =  18
                       spline.jl:33-50   acos, /π, ×NSPL, w 3, 12 mul, 6 add    =  24
  amplitude matrix     rayleigh.jl:21    μ·1.0im                                 =   2
                       mie.jl:77-86      acos, /π, ×(N-1), w 3, 8 mul, 4 add    =  18
                       spline.jl:61-70   (as Mie)                                =  18
  update_E             utils.jl:170-192  frame 77 + ϕ-rotation 12 + S·E (4 cmul
                                         24 + 2 cadd 4) + ψ-rotation 12 + S      = 129 + S
  I_from_E             utils.jl:157-163  3 + 3 + 2 cmul 12 + 2 + 2 + cmul 6 + 2 =  30
  distance             plane.jl:97-105   (x₃ + τ)/(-k₃)                          =   2
                       sphere.jl:69-73   k·x 5, x·x 5, kx² 1, τ² 1, 2 add, √, 1  =  16

  RT peel-off eval     cpu.jl:13-26      direction 2 + update_I + distance
                                         + exp + 8 mul + 4 add + acc 1
  CB bin eval (n > 1)  cpu.jl:85-126     direction 19 + 2 update_E + ΔΦ 12
                                         + τa 16|2 + half-exp 2 + 4 + I_from_E
                                         + τb + 1 + 3 + 4 + I_from_E + e^{iΔΦ} 7
                                         + 16 + I_from_E + 6 + 2 norm_to_ray 34
                                         + ψ 6 + 2ψ 5 + rotate 6 + 4
  CB bin eval (n = 1)  cpu.jl:133-155    direction 19 + update_I + τa + exp 1
                                         + 4 + 34 + 6 + 5 + 6 + 4

Per-event transport (scatter, free path: ~300 flops) is < 0.1% of the bin
work at these workloads and is not counted.  A plane whose layers hold
different scatterers is counted with layer 1's.
"""
from __future__ import annotations

GET_BASIS = 8
NORM_TO_RAY = 17
RAY_TO_NORM = 17
SPH = 10
FRAME = GET_BASIS + NORM_TO_RAY + SPH + RAY_TO_NORM + (2 + RAY_TO_NORM) + 6  # 77
ROTATES = 22
DENSE = 28
PHASE = {"rayleigh": 6, "mie": 18, "spline": 24}
AMPL = {"rayleigh": 2, "mie": 18, "spline": 18}
I_FROM_E = 30
DIST = {"plane": 2, "sphere": 16}


def update_I(sca: str) -> int:
    return FRAME + ROTATES + DENSE + PHASE[sca]


def update_E(sca: str) -> int:
    return FRAME + 12 + 28 + 12 + AMPL[sca]


def f_rt(geom: str, sca: str) -> int:
    """Flops of one peel-off evaluation (cpu.jl:13-26)."""
    return 2 + update_I(sca) + DIST[geom] + 1 + 8 + 4 + 1


def f_cb(geom: str, sca: str) -> int:
    """Flops of one CB bin evaluation with n > 1 (cpu.jl:85-126)."""
    d = DIST[geom]
    return (19 + 2 * update_E(sca) + 12 + d + 2 + 4 + I_FROM_E + d + 1 + 3 + 4 + I_FROM_E + 7 + 16
            + I_FROM_E + 6 + 34 + 6 + 5 + 6 + 4)


def f_cb1(geom: str, sca: str) -> int:
    """Flops of one CB bin evaluation with n = 1 (cpu.jl:133-155)."""
    return 19 + update_I(sca) + DIST[geom] + 1 + 4 + 34 + 6 + 5 + 6 + 4


def kinds(g) -> tuple[str, str]:
    from .geometries import PlaneGeometry
    from .scatterers import RayleighScatterer, SplineScatterer
    geom = "plane" if isinstance(g, PlaneGeometry) else "sphere"
    sca = g.layers[0].scatterer if geom == "plane" else g.scatterer
    if isinstance(sca, RayleighScatterer):
        return geom, "rayleigh"
    return geom, ("spline" if isinstance(sca, SplineScatterer) else "mie")


def algorithmic_flops(g, stats: dict) -> float:
    """Σ count × flops for one solve, from the kernel's exact counters."""
    geom, sca = kinds(g)
    cb1 = stats.get("cb_evals_single", 0)
    return (stats["rt_evals"] * f_rt(geom, sca) + (stats["cb_evals"] - cb1) * f_cb(geom, sca)
            + cb1 * f_cb1(geom, sca))
