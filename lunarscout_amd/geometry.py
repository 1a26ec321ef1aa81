"""Raster grid, projection parameters and float64 host geometry.

Mirrors the reference data contract (src/lunarscout/_numba_horizon/geometry.py:24-116
``ProjectionParameters`` / ``DemGrid``, contract.py:19-38 dtypes and field order).
The heavy float64 setup -- ray origins, directions, chord sampling and fitting --
runs on the GPU (``csrc/lsh_segments.hip``); this module holds only the small
host-side pieces callers need to describe a DEM.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import numpy.typing as npt

from .errors import ContractValidationError

MOON_RADIUS_M = 1_737_400.0

SEGMENT_FIELDS = (
    "start_pixel_x", "start_pixel_y", "x0", "y0", "a1", "a2", "a3", "a4",
    "b1", "b2", "b3", "b4", "s_start_km", "s_end_km",
    "s_start_chord_km", "planar_to_chord_c1", "planar_to_chord_c2",
    "planar_to_chord_c3",
)


@dataclass(frozen=True, slots=True)
class ProjectionParameters:
    """Spherical stereographic projection (geometry.py:24-38)."""

    radius_m: float
    latitude_origin_rad: float
    longitude_origin_rad: float
    scale: float
    false_easting_m: float
    false_northing_m: float

    @classmethod
    def from_array(cls, values: npt.ArrayLike) -> "ProjectionParameters":
        array = np.asarray(values, dtype=np.float64)
        if array.shape != (6,) or not np.all(np.isfinite(array)):
            raise ContractValidationError("projection parameters must be six finite values")
        return cls(*(float(v) for v in array))

    @classmethod
    def south_polar(cls, radius_m: float = MOON_RADIUS_M) -> "ProjectionParameters":
        """+proj=stere +lat_0=-90 +lon_0=0 +k=1 +R=radius."""
        return cls(radius_m, -np.pi / 2.0, 0.0, 1.0, 0.0, 0.0)

    def as_array(self) -> npt.NDArray[np.float64]:
        return np.array(
            (self.radius_m, self.latitude_origin_rad, self.longitude_origin_rad,
             self.scale, self.false_easting_m, self.false_northing_m),
            dtype=np.float64,
        )


PROJECTION_PARAMETER_FIELDS = (
    "radius_m", "latitude_origin_rad", "longitude_origin_rad", "scale",
    "false_easting_m", "false_northing_m",
)


def projection_array(projection) -> npt.NDArray[np.float64]:
    """float64[6] of a projection in PROJECTION_PARAMETER_FIELDS order (contract.py:28-31).

    Reads the named fields, so the reference's own ``ProjectionParameters``
    (geometry.py:24-38, which has no array method) is accepted as well as this
    module's; a sequence of six numbers is taken as already in that order.
    """
    if all(hasattr(projection, f) for f in PROJECTION_PARAMETER_FIELDS):
        values = [getattr(projection, f) for f in PROJECTION_PARAMETER_FIELDS]
    else:
        values = list(np.asarray(projection, dtype=np.float64).reshape(-1))
    array = np.asarray(values, dtype=np.float64)
    if array.shape != (6,) or not np.all(np.isfinite(array)):
        raise ContractValidationError("projection parameters must be six finite values")
    return array


@dataclass(frozen=True, slots=True)
class DemGrid:
    """float32 elevation raster with an affine transform (geometry.py:41-116)."""

    elevation_m: npt.NDArray[np.float32]
    geo_transform: npt.NDArray[np.float64]
    projection: ProjectionParameters

    def __post_init__(self) -> None:
        e = self.elevation_m
        if (not isinstance(e, np.ndarray) or e.dtype != np.dtype("<f4") or e.ndim != 2
                or not e.flags.c_contiguous):
            raise ContractValidationError("DEM elevation must be C-contiguous float32[y, x]")
        t = self.geo_transform
        if (not isinstance(t, np.ndarray) or t.dtype != np.dtype("<f8") or t.shape != (6,)
                or not t.flags.c_contiguous or not np.all(np.isfinite(t))):
            raise ContractValidationError("DEM geotransform must be six C-contiguous float64 values")
        if abs(t[1] * t[5] - t[2] * t[4]) < 1e-12:
            raise ContractValidationError("DEM geotransform is singular")

    @property
    def height(self) -> int:
        return self.elevation_m.shape[0]

    @property
    def width(self) -> int:
        return self.elevation_m.shape[1]

    @property
    def map_resolution_m(self) -> float:
        t = self.geo_transform
        return float((np.hypot(t[1], t[4]) + np.hypot(t[2], t[5])) * 0.5)

    def pixel_to_crs(self, column: float, row: float) -> tuple[float, float]:
        t = self.geo_transform
        return (float(t[0] + t[1] * column + t[2] * row),
                float(t[3] + t[4] * column + t[5] * row))

    def crs_to_pixel(self, x: float, y: float) -> tuple[float, float]:
        t = self.geo_transform
        det = t[1] * t[5] - t[2] * t[4]
        dx, dy = x - t[0], y - t[3]
        return float((t[5] * dx - t[2] * dy) / det), float((-t[4] * dx + t[1] * dy) / det)


def inverse_stereographic(x: float, y: float, p: ProjectionParameters) -> tuple[float, float]:
    """(latitude, longitude) in radians of a projected point (geometry.py:153-174)."""
    xp, yp = x - p.false_easting_m, y - p.false_northing_m
    rho = float(np.hypot(xp, yp))
    if rho < 1e-9:
        return p.latitude_origin_rad, p.longitude_origin_rad
    c = 2.0 * np.arctan2(rho, 2.0 * p.scale * p.radius_m)
    so, co = np.sin(p.latitude_origin_rad), np.cos(p.latitude_origin_rad)
    lat = np.arcsin(np.cos(c) * so + yp * np.sin(c) * co / rho)
    lon = p.longitude_origin_rad + np.arctan2(
        xp * np.sin(c), rho * co * np.cos(c) - yp * so * np.sin(c))
    return float(lat), float(lon)


def project_stereographic(lat: float, lon: float, p: ProjectionParameters) -> tuple[float, float]:
    """Forward spherical stereographic projection (geometry.py:177-193)."""
    sl, cl = np.sin(lat), np.cos(lat)
    so, co = np.sin(p.latitude_origin_rad), np.cos(p.latitude_origin_rad)
    delta = lon - p.longitude_origin_rad
    den = 1.0 + so * sl + co * cl * np.cos(delta)
    if abs(den) < 1e-10:
        den = 1e-10
    k = 2.0 * p.scale * p.radius_m / den
    return (float(k * cl * np.sin(delta) + p.false_easting_m),
            float(k * (co * sl - so * cl * np.cos(delta)) + p.false_northing_m))


def pixel_frame(dem: DemGrid, row: int, column: int):
    """(rotation[3,3] with columns east/north/up, translation[3]) of one pixel (psr.py:29-47)."""
    x, y = dem.pixel_to_crs(column, row)
    lat, lon = inverse_stereographic(x, y, dem.projection)
    cl, sl, co, so = np.cos(lat), np.sin(lat), np.cos(lon), np.sin(lon)
    up = np.array((cl * co, cl * so, sl))
    east = np.array((-so, co, 0.0))
    north = np.array((-sl * co, -sl * so, cl))
    radius = dem.projection.radius_m + float(dem.elevation_m[row, column])
    rotation = np.ascontiguousarray(np.stack((east, north, up)).T)
    translation = -((radius * up) @ rotation)
    return rotation, translation


def local_direction_vector(dem: DemGrid, row: int, column: int, azimuth_deg: float,
                           elevation_deg: float, distance_m: float) -> npt.NDArray[np.float64]:
    """Moon-ME position seen from a pixel at a given local azimuth/elevation.

    The inverse of the pixel-frame transform, as the reference's lighting tests
    build their vectors (tests/numba_horizon/test_phase6b_psr.py:60-73).
    """
    rotation, translation = pixel_frame(dem, row, column)
    az, el = np.deg2rad(azimuth_deg), np.deg2rad(elevation_deg)
    local = np.array((np.sin(az) * np.cos(el), np.cos(az) * np.cos(el), np.sin(el)))
    return (local * distance_m - translation) @ rotation.T
