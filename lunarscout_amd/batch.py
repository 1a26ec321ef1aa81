"""Batched horizons of many landing-site patches: one persistent launch per batch.

``HorizonBatch`` wraps ``lsh_batch_*``: the sites (each an ordered DEM list,
primary first, and an observer height) are registered once; each ``run``
marches up to ``max_tiles`` tiles of any sites in one persistent launch whose
workgroups drain per-XCD work queues, with the per-batch sequence (tile-list
upload, FP64 segments of every tile, queue reset, march) replayed from a
hipGraph. The results are those of the plain per-tile path bit for bit: the
same per-ray arithmetic, a different schedule of work items.

``autotune`` picks the kernel shape (pixel block and azimuths per workgroup)
and the resident workgroup count on a sample of the workload by timing them;
it never touches the step schedule, which is part of the method. Choices are
cached per device and workload key in a JSON file.
"""

from __future__ import annotations

import ctypes
import json
import os
import time
from dataclasses import dataclass
from pathlib import Path
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import DeviceBuffer, Event, check
from .geometry import DemGrid


class _Site(ctypes.Structure):
    _fields_ = [("dem_ids", ctypes.c_int32 * 4), ("dem_count", ctypes.c_int32),
                ("observer_elevation_m", ctypes.c_float)]


class _Tile(ctypes.Structure):
    _fields_ = [("site", ctypes.c_int32), ("tile_column", ctypes.c_int32),
                ("tile_row", ctypes.c_int32), ("reserved", ctypes.c_int32)]


@dataclass(frozen=True)
class Site:
    dems: tuple[DemGrid, ...]
    observer_elevation_m: float = 0.0


def shapes() -> list[tuple[int, int, int]]:
    """(block_x, block_y, azimuths per workgroup) of every batch kernel shape."""
    lib = _lib.load_library()
    n = ctypes.c_int32()
    check(lib.lsh_batch_shape_count(ctypes.byref(n)), "lsh_batch_shape_count")
    out = []
    for i in range(n.value):
        bx, by, az = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib.lsh_batch_shape(i, ctypes.byref(bx), ctypes.byref(by), ctypes.byref(az)),
              "lsh_batch_shape")
        out.append((bx.value, by.value, az.value))
    return out


class HorizonBatch:
    """Persistent batched march over tiles of several sites (``lsh_batch``)."""

    def __init__(self, session, sites: Sequence[Site], *, max_tiles: int,
                 azimuth_count: int = 1440, maximum_distance_m: float = 1_000_000.0):
        self.session = session
        self.sites = list(sites)
        self.max_tiles = int(max_tiles)
        self.azimuth_count = int(azimuth_count)
        table = (_Site * len(self.sites))()
        for i, site in enumerate(self.sites):
            if not 1 <= len(site.dems) <= 4:
                raise ValueError("a site has one to four DEMs")
            for k, dem in enumerate(site.dems):
                table[i].dem_ids[k] = session.dem_id(dem)
            table[i].dem_count = len(site.dems)
            table[i].observer_elevation_m = float(site.observer_elevation_m)
        handle = ctypes.c_void_p()
        check(session.lib.lsh_batch_create(session.handle, ctypes.addressof(table), len(self.sites),
                                           self.max_tiles, self.azimuth_count,
                                           float(maximum_distance_m), ctypes.byref(handle)),
              "lsh_batch_create")
        self.handle = handle.value

    def run(self, tiles: Sequence[tuple[int, int, int]], output: DeviceBuffer, *, config: int = 0,
            workgroups: int = 0, degrees: bool = False, use_graph: bool = True,
            azimuth_major: bool = False) -> None:
        """March ``tiles`` = [(site, tile_column, tile_row), ...] into ``output``.

        ``output``: device float32 [>= len(tiles), 16384, N], or [.., N, 16384]
        with ``azimuth_major``. Stream ordered.
        """
        n = len(tiles)
        if output.nbytes < n * 16384 * self.azimuth_count * 4:
            raise ValueError("output is too small for the batch")
        table = (_Tile * n)()
        for i, (site, column, row) in enumerate(tiles):
            table[i].site, table[i].tile_column, table[i].tile_row = int(site), int(column), int(row)
        check(self.session.lib.lsh_batch_run(self.handle, ctypes.addressof(table), n, int(config),
                                             int(workgroups),
                                             (_lib.HORIZON_DEGREES if degrees else 0)
                                             | (_lib.HORIZON_AZIMUTH_MAJOR if azimuth_major else 0),
                                             1 if use_graph else 0, output.ptr), "lsh_batch_run")

    def close(self) -> None:
        if self.handle:
            self.session.lib.lsh_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def _cache_path() -> Path:
    root = os.environ.get("LSH_AUTOTUNE_CACHE")
    if root:
        return Path(root)
    return Path(os.path.expanduser("~")) / ".cache" / "lunarscout_amd" / "autotune.json"


def autotune(batch: HorizonBatch, sample: Sequence[tuple[int, int, int]], output: DeviceBuffer, *,
             key: str, configs: Sequence[int] | None = None,
             workgroup_factors: Sequence[float] = (1.0, 0.5), use_cache: bool = True) -> dict:
    """Time each (shape, resident workgroups) on ``sample`` tiles; return the fastest.

    The result {"config", "workgroups", "ms", "trials"} is cached under
    (device name, azimuth count, key).
    """
    session = batch.session
    device = session.device_name()
    cache_key = f"{device}|N={batch.azimuth_count}|{key}"
    path = _cache_path()
    if use_cache and path.is_file():
        try:
            cached = json.loads(path.read_text()).get(cache_key)
            if cached:
                return dict(cached, cached=True)
        except (OSError, ValueError):
            pass
    lib = session.lib
    trials = []
    for config in (configs if configs is not None else range(len(shapes()))):
        bx, by, az = shapes()[config]
        count = ctypes.c_int32()
        check(lib.lsh_batch_resident_workgroups(session.handle, config, ctypes.byref(count)),
              "lsh_batch_resident_workgroups")
        resident = max(8, int(count.value))  # every workgroup resident at the kernel's occupancy
        for factor in workgroup_factors:
            wgs = max(8, int(resident * factor) // 8 * 8)
            e0 = Event(session).record()
            batch.run(sample, output, config=config, workgroups=wgs, use_graph=False)
            e1 = Event(session).record()
            session.synchronize()
            trials.append({"config": config, "shape": [bx, by, az], "workgroups": wgs,
                           "ms": e0.elapsed_ms(e1)})
    best = min(trials, key=lambda t: t["ms"])
    result = {"config": best["config"], "workgroups": best["workgroups"], "ms": best["ms"],
              "trials": trials, "tuned_at": time.strftime("%Y-%m-%dT%H:%M:%S")}
    if use_cache:
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            data = json.loads(path.read_text()) if path.is_file() else {}
            data[cache_key] = result
            path.write_text(json.dumps(data, indent=1))
        except (OSError, ValueError):
            pass
    return dict(result, cached=False)
