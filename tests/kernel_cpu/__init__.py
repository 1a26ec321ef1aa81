"""CPU build of the HIP march core (test infrastructure; see kernel_cpu.cpp)."""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

import oracle

_HERE = Path(__file__).resolve().parent
_LIB = _HERE / "libkernel_cpu.so"
_lib = None


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
        _lib = ctypes.CDLL(str(_LIB))
    return _lib


def horizon_tile(pyramids, segment_values, *, tile_column, tile_row, tile_width=128,
                 tile_height=128, subpatch_size=8, observer_elevation_m=0.0, cap=False,
                 pixels=None, threads=0, mode=0):
    """(slopes float32[P, N], winners int32[P, N], counters int64[3])."""
    keep = oracle._Keep()
    segs = np.ascontiguousarray(segment_values, dtype=np.float32)
    N, D = segs.shape[0], segs.shape[2]
    arr = (oracle.LshoPyramid * D)(*[p.struct(keep) for p in pyramids])
    if pixels is not None:
        pixels = keep.add(np.ascontiguousarray(pixels, dtype=np.int32))
        P = len(pixels)
    else:
        P = tile_width * tile_height
    out = np.empty((P, N), dtype=np.float32)
    win = np.empty((P, N), dtype=np.int32)
    counters = np.zeros(3, dtype=np.int64)
    v = ctypes.c_void_p
    status = lib().kcpu_horizon_tile(
        arr, ctypes.c_int32(D), v(segs.ctypes.data), ctypes.c_int32(N),
        ctypes.c_int32(tile_column), ctypes.c_int32(tile_row), ctypes.c_int32(tile_width),
        ctypes.c_int32(tile_height), ctypes.c_int32(subpatch_size),
        ctypes.c_float(observer_elevation_m), ctypes.c_int32((1 if cap else 0) | mode),
        v(pixels.ctypes.data) if pixels is not None else None, ctypes.c_int32(P),
        v(out.ctypes.data), v(win.ctypes.data), v(counters.ctypes.data), ctypes.c_int32(threads))
    if status:
        raise ValueError("invalid request")
    return out, win, counters


def march_segment(segment, pyramid, *, observer_z, pass_index=0, initial=-np.inf, mode=6,
                  cap=False, capacity=100000):
    """The kernel core's pass over one raw segment: (slope, trace rows [n, 12])."""
    keep = oracle._Keep()
    seg = keep.add(np.ascontiguousarray(segment, dtype=np.float32))
    p = pyramid.struct(keep)
    rows = np.zeros((capacity, 12), dtype=np.float32)
    res = ctypes.c_double()
    v = ctypes.c_void_p
    n = lib().kcpu_march_segment(
        v(seg.ctypes.data), ctypes.byref(p), ctypes.c_double(observer_z),
        ctypes.c_int32(pass_index), ctypes.c_double(initial),
        ctypes.c_int32((1 if cap else 0) | mode), v(rows.ctypes.data), ctypes.c_int32(capacity),
        ctypes.byref(res))
    return res.value, rows[:min(n, capacity)].copy()


def trace_ray(pyramids, segment_values, *, tile_column, tile_row, pixel, azimuth,
              tile_width=128, subpatch_size=8, observer_elevation_m=0.0, capacity=200000,
              mode=0):
    keep = oracle._Keep()
    segs = np.ascontiguousarray(segment_values, dtype=np.float32)
    N, D = segs.shape[0], segs.shape[2]
    arr = (oracle.LshoPyramid * D)(*[p.struct(keep) for p in pyramids])
    rows = np.zeros((capacity, 12), dtype=np.float32)
    res = ctypes.c_float()
    v = ctypes.c_void_p
    n = lib().kcpu_trace_ray(arr, ctypes.c_int32(D), v(segs.ctypes.data), ctypes.c_int32(N),
                             ctypes.c_int32(tile_column), ctypes.c_int32(tile_row),
                             ctypes.c_int32(tile_width), ctypes.c_int32(subpatch_size),
                             ctypes.c_float(observer_elevation_m), ctypes.c_int32(mode),
                             ctypes.c_int32(pixel),
                             ctypes.c_int32(azimuth), v(rows.ctypes.data),
                             ctypes.c_int32(capacity), ctypes.byref(res))
    return res.value, rows[:min(n, capacity)].copy()


def atan2(y, x):
    """The lighting path's float64 arctangent (lsh_math.h) of arrays y, x."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(y)
    v = ctypes.c_void_p
    lib().kcpu_atan2(v(y.ctypes.data), v(x.ctypes.data), ctypes.c_int64(y.size), v(out.ctypes.data))
    return out
