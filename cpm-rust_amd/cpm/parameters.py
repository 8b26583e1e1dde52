"""Configuration-design helpers (python/cpm/parameters.py:1-55): pure
arithmetic for choosing a cell count before building a `CPM`."""

from __future__ import annotations

import math
import warnings

from .warnings import CPMLowResolutionWarning

_MIN_CELL_RADIUS_SITES = 5.0


def _implied_cell_radius(target_volume: float, dimensionality: int) -> float:
    if dimensionality == 2:
        return math.sqrt(target_volume / math.pi)
    if dimensionality == 3:
        return (3.0 * target_volume / (4.0 * math.pi)) ** (1.0 / 3.0)
    raise ValueError(
        f"grid must have length 2 or 3 (dimensionality is inferred from it); "
        f"got length {dimensionality}"
    )


def cells_for_density(grid: tuple[int, ...], target_volume: float, phi: float) -> int:
    """`n = round(phi * N_sites / target_volume)`; warns with
    :class:`CPMLowResolutionWarning` when the implied cell radius is below
    ~5 lattice sites."""
    n_sites = math.prod(grid)
    radius = _implied_cell_radius(target_volume, len(grid))
    if radius < _MIN_CELL_RADIUS_SITES:
        warnings.warn(
            f"implied cell radius {radius:.2f} sites (from target_volume="
            f"{target_volume}) is below {_MIN_CELL_RADIUS_SITES:.0f} sites; "
            "lattice discretisation will start dominating the morphology "
            "readouts at this density",
            CPMLowResolutionWarning,
            stacklevel=2,
        )
    return round(phi * n_sites / target_volume)
