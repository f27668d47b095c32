"""Distance metrics accepted as ``distance_metric`` and the geometry each implies
(pair_geometry_for, src/HelperFunctions.jl:390-417; Distances.jl metric types)."""
from __future__ import annotations

from dataclasses import dataclass

from .errors import ArgumentError

GEOM_FLAT, GEOM_SPHERE_DEG, GEOM_SPHERE_RAD = 0, 1, 2


@dataclass(frozen=True)
class Euclidean:
    """Flat space, straight-chord separation."""

    geometry = GEOM_FLAT
    radius = 1.0


@dataclass(frozen=True)
class Haversine:
    """Sphere of ``radius``; coordinates (lon, lat) in degrees (Distances.Haversine).
    Distances.jl's default radius is the Earth's mean radius in metres."""

    radius: float = 6_371_000.0
    geometry = GEOM_SPHERE_DEG


@dataclass(frozen=True)
class SphericalAngle:
    """Unit sphere; coordinates (lon, lat) in radians (Distances.SphericalAngle)."""

    geometry = GEOM_SPHERE_RAD
    radius = 1.0


def geometry_of(metric):
    """(geometry code, radius, coordinate width) or ArgumentError for an unknown metric."""
    if isinstance(metric, Euclidean):
        return GEOM_FLAT, 1.0, None
    if isinstance(metric, Haversine):
        return GEOM_SPHERE_DEG, float(metric.radius), 2
    if isinstance(metric, SphericalAngle):
        return GEOM_SPHERE_RAD, 1.0, 2
    raise ArgumentError(
        f"no pair geometry is defined for distance_metric={type(metric).__name__}. A distance "
        "function fixes which histogram bin a pair falls in, but it does not define the separation "
        "direction or how to compare velocities at two different points, so this package will not "
        "guess one.")
