"""mcrt — MI355X (gfx950) per-ray Monte Carlo RT/CB transport behind the
reference's API (MonteCarloRadiativeTransfer.jl: Config, Layer,
PlaneGeometry, SphereGeometry, RayleighScatterer, MieScatterer,
SplineScatterer, simulate, run_rt).  The transport runs in libmcrt_hip.so; this package marshals the
reference's objects into its C ABI (include/mcrt.h).
"""
from .api import (Context, Plan, ResidentTables, get_context, mueller_plane, mueller_sphere, run_rt,
                  simulate, table_cache_stats)
from .config import DEFAULT_THEB, Config, midpoint_polar_grid
from .geometries import (ArgumentError, Layer, PlaneGeometry, SphereGeometry, gausslegendre,
                         initial_weights)
from . import opcount, workloads  # noqa: E402
from .scatterers import (MieScatterer, RayleighScatterer, SplineScatterer, bhmie, calculate_csrn,
                         mean_free_path, single_scattering_albedo)

__all__ = [
    "Config", "DEFAULT_THEB", "Layer", "PlaneGeometry", "SphereGeometry", "RayleighScatterer",
    "MieScatterer", "SplineScatterer", "simulate", "run_rt", "Plan", "Context", "get_context", "bhmie",
    "calculate_csrn", "gausslegendre", "initial_weights", "midpoint_polar_grid", "mean_free_path",
    "single_scattering_albedo", "ArgumentError", "mueller_plane", "mueller_sphere",
    "ResidentTables", "table_cache_stats",
]
