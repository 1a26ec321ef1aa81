"""CPU oracle of the lunarscout horizon and lighting path.

TEST INFRASTRUCTURE. This package wraps ``liblsh_oracle.so`` (built from
``lsh_oracle.cpp`` by ``oracle/Makefile``), a C++ restatement of the reference
algorithm. Only the tests, the smoke check and the CPU-baseline leg of the
benchmark may import it, and only as the checker or the timed CPU baseline --
the product package never imports it.

Functions mirror the reference (paths under src/lunarscout/_numba_horizon/):

* ``pyramid_layout`` / ``build_pyramid``   pyramid.py:19-31, 110-138
* ``build_samples`` / ``fit_segment``      geometry_numba.py:142-297
* ``segments``                             geometry.py:575-673
* ``march_ray``                            cuda_backend.py:656-874 (+ trace rows of
                                           hierarchy.py:92-276)
* ``horizon_tile`` / ``horizon_tile_f64``  cuda_backend.py:1007-1134, 1445-1552
* ``sun_fraction``                         lightmap.py:32-73
* ``azimuth_elevation``                    psr.py:29-58
* ``lighting_patch``                       lightmap.py:116-215
* ``psr_patch`` / ``psr_reduce``           psr.py:61-140
* ``cbin_encode`` / ``cbin_decode``        file_format.py:39-73, 135-182

Parity status of each function (what pins it) is recorded beside the tests in
``tests/test_oracle_fixtures.py``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liblsh_oracle.so"

FLOAT32 = 1
CAP = 2
REFERENCE_SLOPE = 4

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_vp = ctypes.c_void_p


class LshoDem(ctypes.Structure):
    _fields_ = [
        ("elevation", _vp),
        ("width", _i32),
        ("height", _i32),
        ("geo_transform", _f64 * 6),
        ("projection", _f64 * 6),
    ]


class LshoPyramid(ctypes.Structure):
    _fields_ = [
        ("dem", LshoDem),
        ("mips", _vp),
        ("levels", _vp),
        ("level_count", _i32),
    ]


def build() -> Path:
    """Compile the oracle library in place (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(_LIB_PATH))
    return _lib


def _ptr(array: np.ndarray | None) -> int | None:
    return None if array is None else array.ctypes.data


def _projection_tuple(projection) -> tuple[float, ...]:
    if hasattr(projection, "as_array"):
        return tuple(float(v) for v in projection.as_array())
    if hasattr(projection, "radius_m"):
        return (
            projection.radius_m, projection.latitude_origin_rad,
            projection.longitude_origin_rad, projection.scale,
            projection.false_easting_m, projection.false_northing_m,
        )
    return tuple(float(v) for v in projection)


class _Keep:
    """Keeps numpy buffers alive while ctypes structures point into them."""

    def __init__(self):
        self.items = []

    def add(self, array):
        self.items.append(array)
        return array


def make_dem(dem, keep: _Keep) -> LshoDem:
    elevation = keep.add(np.ascontiguousarray(dem.elevation_m, dtype=np.float32))
    out = LshoDem()
    out.elevation = elevation.ctypes.data
    out.height, out.width = elevation.shape
    for index, value in enumerate(np.asarray(dem.geo_transform, dtype=np.float64)):
        out.geo_transform[index] = float(value)
    for index, value in enumerate(_projection_tuple(dem.projection)):
        out.projection[index] = float(value)
    return out


def pyramid_layout(width: int, height: int) -> tuple[np.ndarray, int]:
    levels = np.zeros((64, 4), dtype=np.int32)
    count = _i32()
    mips = _i64()
    status = lib().lsho_pyramid_layout(
        _i32(width), _i32(height), _vp(levels.ctypes.data), _i32(64),
        ctypes.byref(count), ctypes.byref(mips),
    )
    if status:
        raise ValueError("invalid pyramid dimensions")
    return levels[: count.value].copy(), int(mips.value)


def build_pyramid(elevation: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Return (levels int32[L,4], mips float32[...]) of the max pyramid."""
    elevation = np.ascontiguousarray(elevation, dtype=np.float32)
    height, width = elevation.shape
    levels, count = pyramid_layout(width, height)
    mips = np.empty(count, dtype=np.float32)
    lib().lsho_pyramid_build(
        _vp(elevation.ctypes.data), _i32(width), _i32(height), _vp(mips.ctypes.data)
    )
    return levels, mips


class OraclePyramid:
    """A DEM with its pyramid, in the layout the march oracle reads."""

    def __init__(self, dem, levels: np.ndarray | None = None, mips: np.ndarray | None = None):
        self.dem = dem
        if levels is None or mips is None:
            levels, mips = build_pyramid(dem.elevation_m)
        self.levels = np.ascontiguousarray(levels, dtype=np.int32)
        self.mips = np.ascontiguousarray(mips, dtype=np.float32)
        if self.mips.size == 0:
            self.mips = np.zeros(1, dtype=np.float32)

    def struct(self, keep: _Keep) -> LshoPyramid:
        out = LshoPyramid()
        out.dem = make_dem(self.dem, keep)
        out.mips = keep.add(self.mips).ctypes.data
        out.levels = keep.add(self.levels).ctypes.data
        out.level_count = len(self.levels)
        return out


def build_samples(dem, observer, direction, start_m: float, maximum_m: float) -> np.ndarray:
    keep = _Keep()
    d = make_dem(dem, keep)
    observer = np.ascontiguousarray(observer, dtype=np.float64)
    direction = np.ascontiguousarray(direction, dtype=np.float64)
    samples = np.zeros((16, 8), dtype=np.float64)
    count = _i32()
    lib().lsho_build_samples(
        ctypes.byref(d), _vp(observer.ctypes.data), _vp(direction.ctypes.data),
        _f64(start_m), _f64(maximum_m), _vp(samples.ctypes.data), ctypes.byref(count),
    )
    return samples[: count.value].copy()


def fit_segment(samples: np.ndarray, map_resolution_m: float, observer, radius_m: float,
                fallback_start_m: float) -> np.ndarray:
    padded = np.zeros((16, 8), dtype=np.float64)
    padded[: len(samples)] = samples
    observer = np.ascontiguousarray(observer, dtype=np.float64)
    out = np.zeros(18, dtype=np.float32)
    lib().lsho_fit_segment(
        _vp(padded.ctypes.data), _i32(len(samples)), _f64(map_resolution_m),
        _vp(observer.ctypes.data), _f64(radius_m), _f64(fallback_start_m),
        _vp(out.ctypes.data),
    )
    return out


def segments(dems, *, tile_column: int, tile_row: int, tile_width: int, azimuth_count: int,
             maximum_distance_m: float = 1_000_000.0, observer_elevation_m: float = 0.0,
             subpatch_size: int = 8, threads: int = 0, exact_chain: bool = False) -> np.ndarray:
    """float32[N, C, D, 18] segment tensor of one tile (geometry.py:575-673).

    ``exact_chain`` starts each later DEM at the previous DEM's last float64
    sample distance (the C# / Python-cache chaining) instead of the production
    float32 ``s_end_km``.
    """
    keep = _Keep()
    dems = list(dems)
    array = (LshoDem * len(dems))(*[make_dem(d, keep) for d in dems])
    count = tile_width // subpatch_size + 2
    out = np.zeros((azimuth_count, count * count, len(dems), 18), dtype=np.float32)
    status = lib().lsho_segments(
        array, _i32(len(dems)), _i32(tile_column), _i32(tile_row), _i32(tile_width),
        _i32(azimuth_count), _f64(maximum_distance_m), _f64(observer_elevation_m),
        _i32(subpatch_size), _i32(1 if exact_chain else 0), _vp(out.ctypes.data),
        _i32(threads),
    )
    if status:
        raise ValueError("invalid segment request")
    return out


def march_ray(segment, pyramid: OraclePyramid, *, observer_z: float, radius: float,
              map_resolution: float, pass_index: int, initial: float = -np.inf,
              mode: int = 0, trace_capacity: int = 0):
    """hierarchy_maximum on one interpolated segment; returns (slope, trace rows)."""
    keep = _Keep()
    seg = np.ascontiguousarray(segment, dtype=np.float32)
    p = pyramid.struct(keep)
    trace = np.zeros((max(trace_capacity, 1), 12), dtype=np.float32)
    count = _i32()
    result = _f64()
    lib().lsho_march_ray(
        _vp(seg.ctypes.data), ctypes.byref(p), _f64(observer_z), _f64(radius),
        _f64(map_resolution), _i32(pass_index), _f64(initial), _i32(mode),
        _vp(trace.ctypes.data) if trace_capacity else None, _i32(trace_capacity),
        ctypes.byref(count), ctypes.byref(result),
    )
    rows = min(count.value, trace_capacity)
    return result.value, trace[:rows].copy()


def horizon_tile(pyramids, segment_values: np.ndarray, *, tile_column: int, tile_row: int,
                 tile_width: int = 128, tile_height: int = 128, subpatch_size: int = 8,
                 observer_elevation_m: float = 0.0, mode: int = 0, pixels=None,
                 threads: int = 0, float64: bool = False):
    """Slopes [P, N] of one tile over every DEM pass, plus the traversal counters.

    ``float64=True`` returns the unrounded float64 slopes of method mode.
    ``pixels`` (optional) restricts the computation to those pixel indices.
    """
    keep = _Keep()
    segs = np.ascontiguousarray(segment_values, dtype=np.float32)
    N = segs.shape[0]
    D = segs.shape[2]
    if len(pyramids) != D:
        raise ValueError("pyramid count must match the segment DEM axis")
    arr = (LshoPyramid * D)(*[p.struct(keep) for p in pyramids])
    if pixels is not None:
        pixels = keep.add(np.ascontiguousarray(pixels, dtype=np.int32))
        P = len(pixels)
    else:
        P = tile_width * tile_height
    counters = np.zeros(3, dtype=np.int64)
    if float64:
        out = np.empty((P, N), dtype=np.float64)
        status = lib().lsho_horizon_tile_f64(
            arr, _i32(D), _vp(segs.ctypes.data), _i32(N), _i32(tile_column), _i32(tile_row),
            _i32(tile_width), _i32(tile_height), _i32(subpatch_size),
            _f64(observer_elevation_m), _vp(_ptr(pixels)) if pixels is not None else None,
            _i32(P), _vp(out.ctypes.data), _vp(counters.ctypes.data), _i32(threads),
        )
    else:
        out = np.empty((P, N), dtype=np.float32)
        status = lib().lsho_horizon_tile(
            arr, _i32(D), _vp(segs.ctypes.data), _i32(N), _i32(tile_column), _i32(tile_row),
            _i32(tile_width), _i32(tile_height), _i32(subpatch_size),
            _f64(observer_elevation_m), _i32(mode),
            _vp(_ptr(pixels)) if pixels is not None else None, _i32(P),
            _vp(out.ctypes.data), _vp(counters.ctypes.data), _i32(threads),
        )
    if status:
        raise ValueError("invalid horizon tile request")
    return out, counters


def horizon_tile_frozen(pyramids, eval_dems, segment_values: np.ndarray, *, tile_column: int,
                        tile_row: int, pixels, observer_elevation_m: float = 0.0,
                        threads: int = 0) -> np.ndarray:
    """Float64 slopes [P, N] with the schedule of ``pyramids`` and the heights of ``eval_dems``.

    The march runs on the base DEMs; every level-0 sample it takes is also
    evaluated on the same-shaped ``eval_dems`` (and their observer terrain), and
    the maximum of those slopes is returned -- the derivative convention of the
    adjoint, whose finite differences this gives.
    """
    keep = _Keep()
    segs = keep.add(np.ascontiguousarray(segment_values, dtype=np.float32))
    N, D = segs.shape[0], segs.shape[2]
    arr = (LshoPyramid * D)(*[p.struct(keep) for p in pyramids])
    grids = [keep.add(np.ascontiguousarray(d.elevation_m, dtype=np.float32)) for d in eval_dems]
    for g, p in zip(grids, pyramids):
        if g.shape != (p.dem.height, p.dem.width):
            raise ValueError("evaluation DEMs must match the march DEMs")
    ptrs = (_vp * D)(*[g.ctypes.data for g in grids])
    px = keep.add(np.ascontiguousarray(pixels, dtype=np.int32))
    out = np.empty((len(px), N), dtype=np.float64)
    status = lib().lsho_horizon_tile_frozen(
        arr, ptrs, _i32(D), _vp(segs.ctypes.data), _i32(N), _i32(tile_column), _i32(tile_row),
        _f64(observer_elevation_m), _vp(px.ctypes.data), _i32(len(px)), _vp(out.ctypes.data),
        _i32(threads))
    if status:
        raise ValueError("invalid frozen-schedule request")
    return out


def sun_fraction(horizon: np.ndarray, azimuth_deg: float, elevation_deg: float,
                 *, float32: bool = True) -> float:
    hz = np.ascontiguousarray(horizon, dtype=np.float32)
    out = _f64()
    lib().lsho_sun_fraction(
        _vp(hz.ctypes.data), _i32(len(hz)), _f64(azimuth_deg), _f64(elevation_deg),
        _i32(1 if float32 else 0), ctypes.byref(out),
    )
    return out.value


def azimuth_elevation(dem, row: int, column: int, vectors: np.ndarray):
    keep = _Keep()
    d = make_dem(dem, keep)
    v = np.ascontiguousarray(vectors, dtype=np.float64).reshape(-1, 3)
    az = np.empty(len(v), dtype=np.float64)
    el = np.empty(len(v), dtype=np.float64)
    lib().lsho_azimuth_elevation(
        ctypes.byref(d), _i32(row), _i32(column), _vp(v.ctypes.data), _i32(len(v)),
        _vp(az.ctypes.data), _vp(el.ctypes.data),
    )
    return az, el


def lighting_patch(dem, horizons: np.ndarray, vectors: np.ndarray, *, tile_y: int, tile_x: int,
                   valid_height: int = 128, valid_width: int = 128,
                   float32_slices: bool = False, threads: int = 0, per_time: bool = True):
    """Float64 lighting of one patch.

    Returns dict with ``fraction``/``bytes``/``margin`` (T, vh, vw) when
    ``per_time`` and ``sun_mean``/``visible_fraction`` (vh, vw).
    """
    keep = _Keep()
    d = make_dem(dem, keep)
    hz = np.ascontiguousarray(horizons, dtype=np.float32)
    N = hz.shape[-1]
    if hz.shape != (128, 128, N):
        raise ValueError("horizons must be (128, 128, N)")
    v = np.ascontiguousarray(vectors, dtype=np.float64).reshape(-1, 3)
    T = len(v)
    vh, vw = valid_height, valid_width
    out = {
        "sun_mean": np.empty((vh, vw), dtype=np.float32),
        "visible_fraction": np.empty((vh, vw), dtype=np.float32),
    }
    if per_time:
        out["fraction"] = np.empty((T, vh, vw), dtype=np.float32)
        out["bytes"] = np.empty((T, vh, vw), dtype=np.uint8)
        out["margin"] = np.empty((T, vh, vw), dtype=np.float32)
    lib().lsho_lighting_patch(
        ctypes.byref(d), _vp(hz.ctypes.data), _i32(N), _vp(v.ctypes.data), _i32(T),
        _i32(tile_y), _i32(tile_x), _i32(vh), _i32(vw), _i32(1 if float32_slices else 0),
        _vp(_ptr(out.get("fraction"))) if per_time else None,
        _vp(_ptr(out.get("bytes"))) if per_time else None,
        _vp(_ptr(out.get("margin"))) if per_time else None,
        _vp(out["sun_mean"].ctypes.data), _vp(out["visible_fraction"].ctypes.data),
        _i32(threads),
    )
    return out


def psr_patch(dem, horizons: np.ndarray, vectors: np.ndarray, *, tile_y: int, tile_x: int,
              valid_height: int = 128, valid_width: int = 128,
              sun_angular_size_deg: float = float(np.float32(0.545)), threads: int = 0):
    """(uint8 PSR mask, float64 max margin) of one patch (psr.py:94-140)."""
    keep = _Keep()
    d = make_dem(dem, keep)
    hz = np.ascontiguousarray(horizons, dtype=np.float32)
    N = hz.shape[-1]
    v = np.ascontiguousarray(vectors, dtype=np.float64).reshape(-1, 3)
    out = np.empty((valid_height, valid_width), dtype=np.uint8)
    margin = np.empty((valid_height, valid_width), dtype=np.float64)
    lib().lsho_psr_patch(
        ctypes.byref(d), _vp(hz.ctypes.data), _i32(N), _vp(v.ctypes.data), _i32(len(v)),
        _i32(tile_y), _i32(tile_x), _i32(valid_height), _i32(valid_width),
        _f64(sun_angular_size_deg), _vp(out.ctypes.data), _vp(margin.ctypes.data),
        _i32(threads),
    )
    return out, margin


def psr_reduce(dem, vectors: np.ndarray, azimuth_count: int = 1440):
    keep = _Keep()
    d = make_dem(dem, keep)
    v = np.ascontiguousarray(vectors, dtype=np.float64).reshape(-1, 3)
    indices = np.empty(5 * azimuth_count, dtype=np.int64)
    count = _i32()
    lib().lsho_psr_reduce(
        ctypes.byref(d), _vp(v.ctypes.data), _i32(len(v)), _i32(azimuth_count),
        _vp(indices.ctypes.data), ctypes.byref(count),
    )
    idx = indices[: count.value].copy()
    return np.ascontiguousarray(v[idx]), idx


def cbin_encode(degrees: np.ndarray):
    deg = np.ascontiguousarray(degrees, dtype=np.float32)
    P, N = deg.shape
    out = np.zeros((P, 2 * N), dtype=np.uint8)
    lengths = np.zeros(P, dtype=np.uint16)
    lib().lsho_cbin_encode(
        _vp(deg.ctypes.data), _i32(P), _i32(N), _vp(out.ctypes.data), _vp(lengths.ctypes.data)
    )
    return out, lengths


def cbin_payload(degrees: np.ndarray) -> bytes:
    """The complete .cbin byte stream (u16 LE length + block per pixel)."""
    encoded, lengths = cbin_encode(degrees)
    parts = []
    for row, length in zip(encoded, lengths):
        parts.append(int(length).to_bytes(2, "little"))
        parts.append(row[: int(length)].tobytes())
    return b"".join(parts)


def cbin_decode(payload: bytes, pixel_count: int = 16384, azimuth_count: int = 1440):
    data = np.frombuffer(payload, dtype=np.uint8)
    out = np.empty((pixel_count, azimuth_count), dtype=np.float32)
    status = lib().lsho_cbin_decode(
        _vp(data.ctypes.data), _i64(len(data)), _i32(pixel_count), _i32(azimuth_count),
        _vp(out.ctypes.data),
    )
    if status:
        raise ValueError("structurally invalid compressed horizon payload")
    return out


def cpu_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def trace_tile_ray(pyramids, segment_values, *, tile_column: int, tile_row: int, pixel: int,
                   azimuth: int, tile_width: int = 128, subpatch_size: int = 8,
                   observer_elevation_m: float = 0.0, capacity: int = 200000):
    """Float64 trace rows of one (pixel, azimuth) ray of a tile over every pass."""
    keep = _Keep()
    segs = np.ascontiguousarray(segment_values, dtype=np.float32)
    N, D = segs.shape[0], segs.shape[2]
    arr = (LshoPyramid * D)(*[p.struct(keep) for p in pyramids])
    rows = np.zeros((capacity, 12), dtype=np.float32)
    result = _f64()
    n = lib().lsho_trace_tile_ray(
        arr, _i32(D), _vp(segs.ctypes.data), _i32(N), _i32(tile_column), _i32(tile_row),
        _i32(tile_width), _i32(subpatch_size), _f64(observer_elevation_m), _i32(pixel),
        _i32(azimuth), _vp(rows.ctypes.data), _i32(capacity), ctypes.byref(result))
    return result.value, rows[: min(n, capacity)].copy()


def _sun_disk(half_angle_deg: float = 0.27):
    """The 16-slice table as float32 and max_photons (lightmap.py:15-29)."""
    angle = float(np.float32(half_angle_deg))
    half = np.array([np.float32(np.sqrt(64.0 - (7.5 - i) ** 2) / 8.0 * angle) for i in range(16)],
                    dtype=np.float32)
    total = np.float32(0.0)
    for v in half:
        total = np.float32(total + v)
    return half, np.float32(2.0) * total


def _smooth_clamp(x, W, s):
    value = np.where(x <= -s, 0.0, np.where(x < s, (x + s) ** 2 / (4 * s),
                     np.where(x <= W - s, x, np.where(x < W + s, W - (W + s - x) ** 2 / (4 * s), W))))
    return value


def smoothed_sun_loss(dem, horizons_deg, vectors, *, tile_y: int, tile_x: int, pixels,
                      half_angle_deg: float = 0.27, smoothing_deg: float = 0.02):
    """Per-pixel sum over vectors of the smoothed 16-slice sun fraction, in float64.

    Restates the loss of the adjoint path: the slice walk of _sun_fraction_reference
    (lightmap.py:32-73) with float32 bin positions, each slice clamp
    min(max(x, 0), 2c) replaced by its C1 rounding over +-smoothing_deg; the
    horizons (degrees, [P, N] rows of the given tile-local pixel indices) enter
    in float64. Pixel frames and az/el come from ``azimuth_elevation``.
    """
    half, max_photons = _sun_disk(half_angle_deg)
    hz = np.asarray(horizons_deg, dtype=np.float64)
    N = hz.shape[-1]
    out = np.zeros(len(pixels))
    for k, p in enumerate(pixels):
        row, column = tile_y + int(p) // 128, tile_x + int(p) % 128
        az, el = azimuth_elevation(dem, row, column, vectors)
        el_f = el.astype(np.float32).astype(np.float64)
        # bins per degree, half a bin and the slice step: x4.0, 0.125, 0.135 at 1440 bins
        per_degree = np.float32(N) / np.float32(360.0)
        half_bucket = np.float32(0.5) / per_degree
        slice_step = np.float32(0.135) * (per_degree * np.float32(0.25))
        left_pos = ((az.astype(np.float32) - np.float32(half_angle_deg)) - half_bucket) * per_degree
        left = np.trunc(left_pos).astype(np.int64)
        frac = (left_pos - left.astype(np.float32)).astype(np.float32)
        left = np.where(left < 0, left + N, np.where(left >= N, left - N, left))
        right = np.where(left + 1 >= N, 0, left + 1)
        lo = left.copy()
        photons = np.zeros(len(az))
        for i in range(16):
            c = float(half[i])
            horizon = frac.astype(np.float64) * (hz[k, right] - hz[k, lo]) + hz[k, lo]
            photons += _smooth_clamp(el_f + c - horizon, 2.0 * c, smoothing_deg)
            frac = (frac + slice_step).astype(np.float32)
            step = frac >= np.float32(1.0)
            lo = np.where(step, right, lo)
            right = np.where(step, np.where(right + 1 >= N, 0, right + 1), right)
            frac = np.where(step, (frac - np.float32(1.0)).astype(np.float32), frac)
        out[k] = float(np.sum(photons / float(max_photons)))
    return out
