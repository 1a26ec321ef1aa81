"""Loaders for the golden data under tests/golden (C#/ILGPU captures).

The arrays come from the reference's own test data (see tests/golden/make_fixtures.py).
"""

from __future__ import annotations

import base64
import functools
import json
from pathlib import Path

import numpy as np

from lunarscout_amd.geometry import DemGrid, ProjectionParameters

GOLDEN = Path(__file__).resolve().parent / "golden"


@functools.lru_cache(maxsize=None)
def rays_metadata() -> dict:
    return json.loads((GOLDEN / "reference_rays.json").read_text(encoding="utf-8"))


@functools.lru_cache(maxsize=None)
def rays_arrays() -> dict:
    with np.load(GOLDEN / "reference_rays.npz", allow_pickle=False) as archive:
        return {name: np.ascontiguousarray(archive[name]) for name in archive.files}


@functools.lru_cache(maxsize=None)
def phase4d() -> dict:
    return json.loads((GOLDEN / "phase4d.json").read_text(encoding="utf-8"))


@functools.lru_cache(maxsize=None)
def lightmap_csharp() -> dict:
    return json.loads((GOLDEN / "lightmap_csharp.json").read_text(encoding="utf-8"))


@functools.lru_cache(maxsize=None)
def psr_csharp() -> dict:
    return json.loads((GOLDEN / "psr_csharp.json").read_text(encoding="utf-8"))


def dem(prefix: str) -> DemGrid:
    """DemGrid of a fixture DEM given its array prefix (e.g. 'flat_east__dem_0')."""
    a = rays_arrays()
    return DemGrid(
        np.ascontiguousarray(a[f"{prefix}__elevation_m"], dtype=np.float32),
        np.ascontiguousarray(a[f"{prefix}__geo_transform"], dtype=np.float64),
        ProjectionParameters.from_array(a[f"{prefix}__pyramid__projection_parameters"]),
    )


def pyramid_levels(prefix: str) -> tuple[np.ndarray, np.ndarray]:
    """(levels int32[L,4], concatenated mips float32) of a captured C# pyramid."""
    a = rays_arrays()
    levels = a[f"{prefix}__pyramid__level_metadata"].astype(np.int32)
    parts = [a[f"{prefix}__pyramid__level_{i}"].ravel() for i in range(1, len(levels))]
    mips = np.concatenate(parts).astype(np.float32) if parts else np.empty(0, np.float32)
    return levels, mips


def samples(prefix: str) -> np.ndarray:
    """float64[n, 8] C# ray samples in RAY_SAMPLE_FIELDS order (geometry.py:18-21)."""
    a = rays_arrays()
    fields = ("distance_m", "pixel_x", "pixel_y", "latitude_rad", "longitude_rad",
              "row", "column", "terrain_height_m")
    return np.column_stack([a[f"{prefix}__sample_{f}"] for f in fields])


def decode_f32(text: str) -> np.ndarray:
    return np.frombuffer(base64.b64decode(text), dtype="<f4")


def trace_from_capture(capture: dict) -> np.ndarray:
    fields = ("s_km", "true_distance_m", "level", "cell_x", "cell_y", "pixel_x", "pixel_y",
              "maximum_elevation_m", "sample_elevation_m", "sample_slope", "advance_km",
              "action")
    return np.asarray(
        [[np.nan if row[f] is None else row[f] for f in fields] for row in capture["trace"]],
        dtype=np.float32,
    )


def interpolate_segments(segments, shifts, scale_ratio: float, tx: float, ty: float):
    """Bilinear blend of four shifted float32 segments (kernel_math.py:105-130 semantics)."""
    values = np.asarray(segments, dtype=np.float32).copy()
    d = np.asarray(shifts, dtype=np.float32) * np.float32(scale_ratio)
    values[:, 0] += d[:, 0]
    values[:, 1] += d[:, 1]
    values[:, 2] += d[:, 0]
    values[:, 3] += d[:, 1]
    tx, ty = np.float32(tx), np.float32(ty)
    top = values[0] + (values[1] - values[0]) * tx
    bottom = values[2] + (values[3] - values[2]) * tx
    return np.ascontiguousarray(top + (bottom - top) * ty, dtype=np.float32)
