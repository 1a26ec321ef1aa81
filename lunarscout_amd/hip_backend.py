"""HIP horizon session: the GPU side of the horizon path.

``HipSession`` is the drop-in for the reference's ``CudaSession``
(src/lunarscout/_numba_horizon/cuda_backend.py:1157-1577) on the methods the
production pipeline calls:

* ``build_max_pyramid(dem)``                      cuda_backend.py:1260-1281
* ``subpatch_hierarchical_all_passes(...)``       cuda_backend.py:1445-1552
* ``build_subpatch_segments(...)``                geometry.py:575-673 (on the GPU)

plus ``horizon_tile`` which keeps segments and slopes on the device between the
two kernels (the path the benchmark times).
"""

from __future__ import annotations

import ctypes
import threading
import time
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import DeviceBuffer, check
from .errors import ContractValidationError
from .geometry import DemGrid, ProjectionParameters, projection_array

PATCH_SIZE = 128
AZIMUTH_COUNT = 1440
SUBPATCH_SIZE = 8
MAXIMUM_DISTANCE_M = 1_000_000.0


@dataclass(frozen=True)
class PyramidArrays:
    """Host copy of a DEM's max pyramid (contract.py:274-344), with its DEM."""

    dem: DemGrid
    levels: np.ndarray  # int32[L, 4]
    mips: np.ndarray    # float32[sum of levels 1..L-1]

    @property
    def level0(self) -> np.ndarray:
        return self.dem.elevation_m


def dem_of_pyramid(pyramid) -> tuple[DemGrid, np.ndarray]:
    """The DEM a pyramid describes and its concatenated mips.

    Takes this module's ``PyramidArrays`` or the reference contract's
    (contract.py:274-281): ``level0`` float32 [H, W], ``mips`` float32, ``levels``
    int32 [L, 4], ``map_parameters`` float32[11] (radius, scale, false
    easting/northing, 1 / determinant, the six transform coefficients;
    pyramid.py:78-107) and ``projection_parameters`` float32[6]. The march reads
    the map resolution and radius in float32 from those values, so rebuilding a
    float64 transform from the float32 parameters changes nothing it computes;
    the level table must be this pyramid layout's.
    """
    if hasattr(pyramid, "dem"):
        return pyramid.dem, np.ascontiguousarray(pyramid.mips, dtype=np.float32)
    try:
        level0 = np.ascontiguousarray(pyramid.level0, dtype=np.float32)
        map_parameters = np.asarray(pyramid.map_parameters, dtype=np.float64).reshape(-1)
        projection = np.asarray(pyramid.projection_parameters, dtype=np.float64).reshape(-1)
        levels = np.asarray(pyramid.levels, dtype=np.int32)
        mips = np.ascontiguousarray(pyramid.mips, dtype=np.float32)
    except AttributeError as exc:
        raise ContractValidationError(
            "pyramids must carry level0, mips, levels, map_parameters and "
            "projection_parameters") from exc
    if level0.ndim != 2 or map_parameters.shape != (11,):
        raise ContractValidationError("pyramid level 0 must be 2-D with 11 map parameters")
    dem = DemGrid(level0, np.ascontiguousarray(map_parameters[5:11]),
                  ProjectionParameters.from_array(projection))
    expected, _ = pyramid_layout(dem.width, dem.height)
    if levels.shape != expected.shape or not np.array_equal(levels, expected):
        raise ContractValidationError("pyramid level metadata does not match the DEM's layout")
    return dem, mips


def pyramid_layout(width: int, height: int) -> tuple[np.ndarray, int]:
    lib = _lib.load_library()
    levels = np.zeros((16, 4), dtype=np.int32)
    count = ctypes.c_int32()
    mips = ctypes.c_int64()
    check(lib.lsh_pyramid_layout(width, height, levels.ctypes.data, 16, ctypes.byref(count),
                                 ctypes.byref(mips)), "lsh_pyramid_layout")
    return levels[: count.value].copy(), int(mips.value)


@dataclass
class _DemEntry:
    dem: DemGrid
    dem_id: int


class HipSession(_lib.Session):
    """An explicitly initialised HIP device session with resident DEMs."""

    def __init__(self, device_id: int = 0):
        super().__init__(device_id)
        self._dems: dict[int, _DemEntry] = {}
        self._dem_lock = threading.Lock()
        self._segments: DeviceBuffer | None = None
        self._output: DeviceBuffer | None = None
        self._pyramid_dems: dict[int, tuple[object, DemGrid]] = {}
        self.last_production_timings: dict | None = None

    def close(self) -> None:
        """Free the session's buffers, its resident DEMs and the stream."""
        for name in ("_segments", "_output"):
            buf = getattr(self, name, None)
            if buf is not None:
                buf.free()
                setattr(self, name, None)
        self._dems = {}
        self._pyramid_dems = {}
        super().close()

    # -- DEM residency ---------------------------------------------------
    def dem_id(self, dem: DemGrid, mips: np.ndarray | None = None) -> int:
        """Register ``dem`` on the device once (pyramid built there unless given)."""
        with self._dem_lock:
            entry = self._dems.get(id(dem))
            if entry is not None and entry.dem is dem:
                return entry.dem_id
            elevation = np.ascontiguousarray(dem.elevation_m, dtype=np.float32)
            gt = np.ascontiguousarray(dem.geo_transform, dtype=np.float64)
            proj = projection_array(dem.projection)
            mips_ptr = None
            if mips is not None:
                _, expected = pyramid_layout(dem.width, dem.height)
                mips = np.ascontiguousarray(mips, dtype=np.float32)
                if mips.size != expected:
                    raise ContractValidationError("pyramid payload length does not match the DEM")
                mips_ptr = mips.ctypes.data
            out = ctypes.c_int32()
            check(self.lib.lsh_dem_create(self.handle, elevation.ctypes.data, dem.width,
                                          dem.height, gt.ctypes.data, proj.ctypes.data, mips_ptr,
                                          ctypes.byref(out)), "lsh_dem_create")
            self._dems[id(dem)] = _DemEntry(dem, int(out.value))
            return int(out.value)

    def forget(self, dem: DemGrid) -> None:
        """Release a registered DEM's device copy (a later use registers it again)."""
        with self._dem_lock:
            entry = self._dems.pop(id(dem), None)
        if entry is not None and entry.dem is dem:
            check(self.lib.lsh_dem_destroy(self.handle, entry.dem_id), "lsh_dem_destroy")

    def build_max_pyramid(self, dem: DemGrid) -> PyramidArrays:
        """Factor-four max pyramid built on the GPU (cuda_backend.py:1260-1281)."""
        dem_id = self.dem_id(dem)
        levels, count = pyramid_layout(dem.width, dem.height)
        mips = np.empty(count, dtype=np.float32)
        if count:
            check(self.lib.lsh_dem_pyramid(self.handle, dem_id, mips.ctypes.data, count),
                  "lsh_dem_pyramid")
        return PyramidArrays(dem, levels, mips)

    def _ids(self, dems: Sequence[DemGrid]) -> np.ndarray:
        if not 1 <= len(dems) <= 4:
            raise ContractValidationError("between one and four DEMs are supported")
        return np.array([self.dem_id(d) for d in dems], dtype=np.int32)

    # -- segments -------------------------------------------------------------
    def segments_device(self, dems: Sequence[DemGrid], *, tile_column: int, tile_row: int,
                        tile_width: int = PATCH_SIZE, azimuth_count: int = AZIMUTH_COUNT,
                        maximum_distance_m: float = MAXIMUM_DISTANCE_M,
                        observer_elevation_m: float = 0.0,
                        subpatch_size: int = SUBPATCH_SIZE,
                        out: DeviceBuffer | None = None) -> DeviceBuffer:
        ids = self._ids(dems)
        count = tile_width // subpatch_size + 2
        shape = (azimuth_count, count * count, len(dems), 18)
        if out is None or out.shape != shape:
            out = DeviceBuffer.empty(self, shape, np.float32)
        check(self.lib.lsh_segments_tile(self.handle, ids.ctypes.data, len(ids), tile_column,
                                         tile_row, tile_width, azimuth_count,
                                         float(maximum_distance_m), float(observer_elevation_m),
                                         subpatch_size, out.ptr), "lsh_segments_tile")
        return out

    def build_subpatch_segments(self, dems: Sequence[DemGrid], **kwargs) -> np.ndarray:
        """float32[N, C, D, 18] segment tensor of one tile, computed on the GPU."""
        buf = self.segments_device(dems, **kwargs)
        try:
            return buf.to_host()
        finally:
            buf.free()

    # -- march ------------------------------------------------------------------
    def march_device(self, dems: Sequence[DemGrid], segments: DeviceBuffer, *, tile_column: int,
                     tile_row: int, tile_width: int = PATCH_SIZE, tile_height: int = PATCH_SIZE,
                     subpatch_size: int = SUBPATCH_SIZE, observer_elevation_m: float = 0.0,
                     flags: int = 0, output: DeviceBuffer | None = None,
                     counters: DeviceBuffer | None = None,
                     azimuth_major: bool = False) -> DeviceBuffer:
        """March one tile into ``output``: float32 [pixels, N], or [N, pixels] azimuth-major."""
        ids = self._ids(dems)
        azimuth_count = segments.shape[0]
        pixels = tile_width * tile_height
        if azimuth_major:
            flags |= _lib.HORIZON_AZIMUTH_MAJOR
        shape = (azimuth_count, pixels) if flags & _lib.HORIZON_AZIMUTH_MAJOR else (pixels, azimuth_count)
        if output is None:
            output = DeviceBuffer.empty(self, shape, np.float32)
        elif output.shape != shape or output.dtype != np.float32:
            raise ContractValidationError(
                f"march output must be float32 {shape}, got {output.dtype} {output.shape}")
        check(self.lib.lsh_horizon_tile(
            self.handle, ids.ctypes.data, len(ids), segments.ptr, azimuth_count, tile_column,
            tile_row, tile_width, tile_height, subpatch_size, float(observer_elevation_m),
            flags, output.ptr, counters.ptr if counters is not None else None),
            "lsh_horizon_tile")
        return output

    def subpatch_hierarchical_all_passes(
        self, segment_values: np.ndarray, pyramids: Sequence[PyramidArrays], *,
        tile_column: int, tile_row: int, tile_width: int, tile_height: int,
        subpatch_size: int, observer_elevation_m: float = 0.0,
    ) -> tuple[np.ndarray, list[float]]:
        """Every DEM pass on one resident slope buffer (cuda_backend.py:1445-1552).

        ``pyramids`` are this module's ``PyramidArrays`` or the reference
        contract's (see ``dem_of_pyramid``), one per DEM pass, primary first.
        Returns float32 slopes [tile_width * tile_height, N] and the kernel seconds
        (all passes run in one launch, so the list holds one entry).
        """
        segments = np.ascontiguousarray(segment_values, dtype=np.float32)
        if segments.ndim != 4 or segments.shape[-1] != 18:
            raise ValueError("segment_values must have shape (azimuth, center, DEM, 18)")
        if len(pyramids) != segments.shape[2]:
            raise ValueError("pyramid count must match the segment DEM axis")
        if segments.shape[1] != (tile_width // subpatch_size + 2) ** 2:
            raise ValueError("segment center count does not match tile/subpatch layout")
        dems = []
        for p in pyramids:
            # one DemGrid per pyramid object, kept with it, so its device copy is reused
            held = self._pyramid_dems.get(id(p))
            if held is None or held[0] is not p:
                dem, mips = dem_of_pyramid(p)
                self.dem_id(dem, mips)
                held = self._pyramid_dems[id(p)] = (p, dem)
            dems.append(held[1])
        started = time.perf_counter()
        seg_buf = DeviceBuffer.from_host(self, segments)
        upload = time.perf_counter() - started
        try:
            kernel_started = time.perf_counter()
            out = self.march_device(dems, seg_buf, tile_column=tile_column, tile_row=tile_row,
                                    tile_width=tile_width, tile_height=tile_height,
                                    subpatch_size=subpatch_size,
                                    observer_elevation_m=observer_elevation_m)
            self.synchronize()
            kernel = time.perf_counter() - kernel_started
            copy_started = time.perf_counter()
            host = out.to_host()
            out.free()
            self.last_production_timings = {
                "segment_upload_seconds": upload,
                "kernel_wall_seconds": kernel,
                "copy_back_seconds": time.perf_counter() - copy_started,
            }
            return host, [kernel]
        finally:
            seg_buf.free()

    def horizon_tile(self, dems: Sequence[DemGrid], *, tile_column: int, tile_row: int,
                     azimuth_count: int = AZIMUTH_COUNT,
                     maximum_distance_m: float = MAXIMUM_DISTANCE_M,
                     observer_elevation_m: float = 0.0, degrees: bool = False,
                     cap: bool = False, counters: bool = False):
        """Segments + march of one 128x128 tile on the GPU; host float32 [16384, N].

        With ``counters`` returns (values, int64[4] level-0 samples / bound tests / steps /
        64 x longest-lane steps per wavefront).
        """
        self._segments = self.segments_device(
            dems, tile_column=tile_column, tile_row=tile_row, azimuth_count=azimuth_count,
            maximum_distance_m=maximum_distance_m, observer_elevation_m=observer_elevation_m,
            out=self._segments)
        flags = (_lib.HORIZON_DEGREES if degrees else 0) | (_lib.HORIZON_CAP if cap else 0)
        counter_buf = None
        if counters:
            counter_buf = DeviceBuffer.empty(self, (4,), np.int64)
            counter_buf.zero()
            flags |= _lib.HORIZON_COUNTERS
        if self._output is not None and self._output.shape != (PATCH_SIZE * PATCH_SIZE, azimuth_count):
            self._output.free()
            self._output = None
        self._output = self.march_device(dems, self._segments, tile_column=tile_column,
                                         tile_row=tile_row,
                                         observer_elevation_m=observer_elevation_m,
                                         flags=flags, output=self._output, counters=counter_buf)
        values = self._output.to_host()
        if counter_buf is not None:
            c = counter_buf.to_host()
            counter_buf.free()
            return values, c
        return values
