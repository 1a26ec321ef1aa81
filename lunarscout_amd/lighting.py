"""HIP lighting session: lightmap, Earth margin and PSR from stored horizons.

``HipLightingSession`` is the drop-in for the reference's ``LightmapCudaSession``
(src/lunarscout/_numba_horizon/lightmap_cuda.py:187-276) and ``PsrCudaSession``
(psr_cuda.py:191-375) on the calls the product pipelines make:

* ``iter_patch_tiles``           uint8 sun-fraction bytes per time (lightmap_cpu.py:233-237)
* ``iter_patch_fraction_tiles``  float32 sun fraction per time      (lightmap_cpu.py:221-225)
* ``iter_patch_margin_tiles``    float32 body margin per time       (lightmap_cpu.py:227-231)
* ``compute_patch``              uint8 PSR mask of one patch         (psr_cuda.py:238-375)

plus ``lighting_tiles``, the streamed many-tile pass (sun fraction, Earth
visibility and PSR together) on device-resident horizons.
"""

from __future__ import annotations

import ctypes
from collections.abc import Iterator
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import DeviceBuffer, check
from .errors import InputError
from .geometry import DemGrid, projection_array

PATCH_SIZE = 128
PATCH_PIXELS = PATCH_SIZE * PATCH_SIZE
SUN_HALF_ANGLE_DEG = 0.27          # lightmap.py:15
SUN_ANGULAR_SIZE_DEG = 0.545       # psr.py:12


def validate_vectors(vectors_m) -> np.ndarray:
    """Finite float64 Moon-ME vectors shaped (time, 3) (psr.py:17-26)."""
    values = np.ascontiguousarray(vectors_m, dtype=np.float64)
    if values.ndim != 2 or values.shape[1:] != (3,) or values.shape[0] == 0 or not np.all(
            np.isfinite(values)):
        raise ValueError("Moon-ME vectors must be finite values shaped (time, 3)")
    return values


def reduce_sun_vectors_for_psr(dem: DemGrid, sun_vectors_m, azimuth_count: int = 1440):
    """Five-viewpoint per-bin maximum-elevation reduction (psr.py:61-91)."""
    vectors = validate_vectors(sun_vectors_m)
    lib = _lib.load_library()
    elevation = np.ascontiguousarray(dem.elevation_m, dtype=np.float32)
    gt = np.ascontiguousarray(dem.geo_transform, dtype=np.float64)
    proj = projection_array(dem.projection)
    indices = np.empty(5 * azimuth_count, dtype=np.int64)
    kept = ctypes.c_int32()
    check(lib.lsh_psr_reduce(elevation.ctypes.data, dem.width, dem.height, gt.ctypes.data,
                             proj.ctypes.data, vectors.ctypes.data, len(vectors), azimuth_count,
                             indices.ctypes.data, ctypes.byref(kept)), "lsh_psr_reduce")
    idx = indices[: kept.value].copy()
    return np.ascontiguousarray(vectors[idx]), idx


@dataclass
class LightingOutputs:
    """Device outputs of ``lighting_tiles`` (None when not requested)."""

    sun_bytes: DeviceBuffer | None = None
    sun_fraction: DeviceBuffer | None = None
    sun_mean: DeviceBuffer | None = None
    earth_margin: DeviceBuffer | None = None
    earth_visible: DeviceBuffer | None = None
    psr: DeviceBuffer | None = None
    psr_margin: DeviceBuffer | None = None

    def free(self) -> None:
        for value in self.__dict__.values():
            if value is not None:
                value.free()


def _tile_array(tiles) -> np.ndarray:
    arr = np.ascontiguousarray(np.asarray(tiles, dtype=np.int32).reshape(-1, 2))
    return arr  # (tile_column, tile_row) pairs == lsh_tile


def _check_device_horizons(horizons, tile_count: int, azimuth_major: bool) -> int:
    """The azimuth count of a device horizon buffer holding ``tile_count`` tiles.

    The buffer must be float32 [tile][16384][N] (or [tile][N][16384]
    azimuth-major) with at least ``tile_count`` tiles, so no kernel reads past it.
    """
    if not isinstance(horizons, DeviceBuffer):
        raise InputError("horizons must be a device buffer", code="lighting_input_invalid")
    if horizons.dtype != np.float32 or len(horizons.shape) < 2:
        raise InputError("horizons must be float32 [tile][pixel][azimuth] on the device",
                         code="lighting_input_invalid", details={"dtype": str(horizons.dtype)})
    N = int(horizons.shape[-2] if azimuth_major else horizons.shape[-1])
    pixels = int(horizons.shape[-1] if azimuth_major else horizons.shape[-2])
    if pixels != PATCH_PIXELS or N < 2:
        raise InputError("horizon tiles must hold 16384 pixels and at least 2 azimuths",
                         code="lighting_input_invalid", details={"shape": list(horizons.shape)})
    if horizons.nbytes < tile_count * PATCH_PIXELS * N * 4:
        raise InputError("the horizon buffer holds fewer tiles than requested",
                         code="lighting_input_invalid",
                         details={"tiles": tile_count, "shape": list(horizons.shape)})
    return N


def _check_device_vectors(buf, name: str, minimum: int = 0) -> int:
    """The vector count of a device [T][3] float64 buffer (at least ``minimum``)."""
    if buf.dtype != np.float64 or len(buf.shape) != 2 or buf.shape[1] != 3:
        raise InputError(f"{name} vectors must be float64 [T][3] on the device",
                         code="lighting_input_invalid", details={"shape": list(buf.shape)})
    if buf.nbytes < buf.shape[0] * 24 or buf.shape[0] < minimum:
        raise InputError(f"{name} vectors hold fewer entries than required",
                         code="lighting_input_invalid",
                         details={"have": int(buf.shape[0]), "need": int(minimum)})
    return int(buf.shape[0])


class HipLightingSession:
    """Lightmap / margin / PSR on the GPU with fixed-shape reusable buffers."""

    def __init__(self, session=None, *, device_id: int = 0, time_batch_size: int = 32,
                 sun_half_angle_deg: float = SUN_HALF_ANGLE_DEG,
                 sun_angular_size_deg: float = SUN_ANGULAR_SIZE_DEG):
        from .hip_backend import HipSession

        if time_batch_size < 1:
            raise ValueError("time_batch_size must be positive")
        self.session = session if session is not None else HipSession(device_id)
        self.time_batch_size = int(time_batch_size)
        self.sun_half_angle_deg = float(sun_half_angle_deg)
        self.sun_angular_size_deg = float(sun_angular_size_deg)
        self._horizon_buf: DeviceBuffer | None = None
        self._vector_cache: dict[bytes, DeviceBuffer] = {}

    # -- bulk pass ----------------------------------------------------------------
    def lighting_tiles(self, dem: DemGrid, tiles, horizons: DeviceBuffer, *,
                       sun=None, earth=None, psr_vectors=None, azimuth_major: bool = False,
                       sun_bytes: bool = False, sun_fraction: bool = False,
                       sun_mean: bool = True, earth_margin: bool = False,
                       earth_visible: bool = True, psr: bool = True, psr_margin: bool = False,
                       outputs: LightingOutputs | None = None) -> LightingOutputs:
        """One launch over many tiles; horizons already on the device (degrees)."""
        s = self.session
        dem_id = s.dem_id(dem)
        tile_arr = _tile_array(tiles)
        n = len(tile_arr)
        N = _check_device_horizons(horizons, n, azimuth_major)
        bufs = {}
        counts = {}
        for name, vec in (("sun", sun), ("earth", earth), ("psr", psr_vectors)):
            if vec is None:
                bufs[name] = None
                counts[name] = 0
            elif isinstance(vec, DeviceBuffer):
                bufs[name] = vec
                counts[name] = _check_device_vectors(vec, name)
            else:
                v = validate_vectors(vec)
                bufs[name] = DeviceBuffer.from_host(s, v)
                counts[name] = len(v)
        out = outputs or LightingOutputs()
        T, E = counts["sun"], counts["earth"]
        want = {
            "sun_bytes": (sun_bytes and T, (n, T, PATCH_PIXELS), np.uint8, _lib.LIGHT_SUN_BYTES),
            "sun_fraction": (sun_fraction and T, (n, T, PATCH_PIXELS), np.float32, _lib.LIGHT_SUN_FRACTION),
            "sun_mean": (sun_mean and T, (n, PATCH_PIXELS), np.float32, _lib.LIGHT_SUN_MEAN),
            "earth_margin": (earth_margin and E, (n, E, PATCH_PIXELS), np.float32, _lib.LIGHT_EARTH_MARGIN),
            "earth_visible": (earth_visible and E, (n, PATCH_PIXELS), np.float32, _lib.LIGHT_EARTH_VISIBLE),
            "psr": (psr and counts["psr"], (n, PATCH_PIXELS), np.uint8, _lib.LIGHT_PSR),
            "psr_margin": (psr_margin and counts["psr"], (n, PATCH_PIXELS), np.float32, _lib.LIGHT_PSR_MARGIN),
        }
        flags = 0
        ptrs = {}
        for name, (enabled, shape, dtype, flag) in want.items():
            if not enabled:
                ptrs[name] = None
                continue
            buf = getattr(out, name)
            if buf is None or buf.shape != shape:
                buf = DeviceBuffer.empty(s, shape, dtype)
                setattr(out, name, buf)
            ptrs[name] = buf.ptr
            flags |= flag
        check(s.lib.lsh_lighting_tiles(
            s.handle, dem_id, tile_arr.ctypes.data, n, horizons.ptr, N, 1 if azimuth_major else 0,
            bufs["sun"].ptr if bufs["sun"] else None, T,
            bufs["earth"].ptr if bufs["earth"] else None, E,
            bufs["psr"].ptr if bufs["psr"] else None, counts["psr"],
            self.sun_half_angle_deg, self.sun_angular_size_deg, flags,
            ptrs["sun_bytes"], ptrs["sun_fraction"], ptrs["sun_mean"], ptrs["earth_margin"],
            ptrs["earth_visible"], ptrs["psr"], ptrs["psr_margin"]), "lsh_lighting_tiles")
        s.synchronize()
        for name, vec in (("sun", sun), ("earth", earth), ("psr", psr_vectors)):
            if bufs[name] is not None and not isinstance(vec, DeviceBuffer):
                bufs[name].free()
        return out

    # -- duration reductions -------------------------------------------------------
    def duration_tiles(self, dem: DemGrid, tiles, horizons: DeviceBuffer, plan, sun, earth=None, *,
                       azimuth_major: bool = False, out: DeviceBuffer | None = None) -> DeviceBuffer:
        """Longest-duration / safe-haven bands of many tiles in one launch
        (``lsh_duration_tiles``): float32 [bands][tile][16384] on the device.
        ``plan`` is a ``durations.DurationPlan``; sun / earth are host arrays or
        device buffers of [T][3] Moon-ME positions."""
        s = self.session
        dem_id = s.dem_id(dem)
        tile_arr = _tile_array(tiles)
        n = len(tile_arr)
        N = _check_device_horizons(horizons, n, azimuth_major)
        spec = plan.spec()
        owned = []

        def device(vec, name):
            if vec is None:
                return vec
            if isinstance(vec, DeviceBuffer):
                _check_device_vectors(vec, name, plan.time_count)
                return vec
            host = validate_vectors(vec)
            if len(host) < plan.time_count:
                raise InputError(f"{name} vectors hold fewer entries than the plan's times",
                                 code="lighting_input_invalid",
                                 details={"have": len(host), "need": plan.time_count})
            buf = DeviceBuffer.from_host(s, host)
            owned.append(buf)
            return buf

        if plan.needs_earth and earth is None:
            raise InputError("this duration plan needs Earth vectors", code="lighting_input_invalid")
        sun_buf = device(sun, "sun")
        earth_buf = device(earth if plan.needs_earth else None, "earth")
        need = ctypes.c_int64()
        check(s.lib.lsh_duration_workspace_bytes(ctypes.byref(spec), n, ctypes.byref(need)),
              "lsh_duration_workspace_bytes")
        workspace = DeviceBuffer.empty(s, (max(int(need.value), 8),), np.uint8)
        owned.append(workspace)
        shape = (plan.bands, n, PATCH_PIXELS)
        if out is None or out.shape != shape:
            out = DeviceBuffer.empty(s, shape, np.float32)
        try:
            check(s.lib.lsh_duration_tiles(
                s.handle, dem_id, tile_arr.ctypes.data, n, horizons.ptr, N, 1 if azimuth_major else 0,
                sun_buf.ptr, earth_buf.ptr if earth_buf is not None else None,
                self.sun_half_angle_deg, ctypes.byref(spec), workspace.ptr, out.ptr),
                "lsh_duration_tiles")
            s.synchronize()
        finally:
            for buf in owned:
                buf.free()
        return out

    def duration_patch(self, dem, horizons_deg, plan, sun_vectors_m, earth_vectors_m=None, *,
                       tile_y, tile_x, valid_height=PATCH_SIZE, valid_width=PATCH_SIZE):
        """One patch's bands, float32 [valid_height, valid_width] each."""
        self._check_patch(dem, tile_y, tile_x, valid_height, valid_width)
        hz = self._upload_patch(_finite_patch(horizons_deg).reshape(PATCH_SIZE, PATCH_SIZE, -1))
        out = self.duration_tiles(dem, [(tile_x, tile_y)], hz, plan, sun_vectors_m, earth_vectors_m)
        values = out.to_host().reshape(plan.bands, PATCH_SIZE, PATCH_SIZE)
        out.free()
        return tuple(np.ascontiguousarray(b[:valid_height, :valid_width]) for b in values)

    # -- per-patch drop-ins -------------------------------------------------------
    def _upload_patch(self, horizons_deg) -> DeviceBuffer:
        horizons = np.ascontiguousarray(horizons_deg, dtype=np.float32)
        if horizons.ndim != 3 or horizons.shape[:2] != (PATCH_SIZE, PATCH_SIZE):
            raise ValueError("horizons must have shape (128, 128, N)")
        if not np.all(np.isfinite(horizons)):
            raise ValueError("horizons must contain finite elevations")
        shape = (1, PATCH_PIXELS, horizons.shape[2])
        if self._horizon_buf is None or self._horizon_buf.shape != shape:
            if self._horizon_buf is not None:
                self._horizon_buf.free()
            self._horizon_buf = DeviceBuffer.empty(self.session, shape, np.float32)
        self._horizon_buf.upload(horizons.reshape(shape))
        return self._horizon_buf

    @staticmethod
    def _check_patch(dem: DemGrid, tile_y: int, tile_x: int, valid_height: int, valid_width: int):
        if not 1 <= valid_width <= PATCH_SIZE or not 1 <= valid_height <= PATCH_SIZE:
            raise ValueError("valid patch dimensions must be between 1 and 128")
        if tile_x < 0 or tile_y < 0 or tile_x + valid_width > dem.width or tile_y + valid_height > dem.height:
            raise ValueError("patch falls outside the DEM")

    def _signal_batches(self, dem, horizons_deg, vectors_m, *, tile_y, tile_x, valid_height,
                        valid_width, margin: bool):
        self._check_patch(dem, tile_y, tile_x, valid_height, valid_width)
        vectors = validate_vectors(vectors_m)
        hz = self._upload_patch(horizons_deg)
        out = LightingOutputs()
        try:
            for start in range(0, len(vectors), self.time_batch_size):
                batch = vectors[start:start + self.time_batch_size]
                if margin:
                    self.lighting_tiles(dem, [(tile_x, tile_y)], hz, earth=batch, sun_mean=False,
                                        earth_visible=False, psr=False, earth_margin=True,
                                        outputs=out)
                    values = out.earth_margin.to_host()
                else:
                    self.lighting_tiles(dem, [(tile_x, tile_y)], hz, sun=batch, sun_mean=False,
                                        earth_visible=False, psr=False, sun_fraction=True,
                                        outputs=out)
                    values = out.sun_fraction.to_host()
                grid = values.reshape(len(batch), PATCH_SIZE, PATCH_SIZE)
                yield grid[:, :valid_height, :valid_width]
        finally:
            out.free()

    def iter_patch_fraction_tiles(self, dem, horizons_deg, sun_vectors_m, *, tile_y, tile_x,
                                  valid_height=PATCH_SIZE, valid_width=PATCH_SIZE
                                  ) -> Iterator[np.ndarray]:
        for batch in self._signal_batches(dem, horizons_deg, sun_vectors_m, tile_y=tile_y,
                                          tile_x=tile_x, valid_height=valid_height,
                                          valid_width=valid_width, margin=False):
            for tile in batch:
                yield np.ascontiguousarray(tile)

    def iter_patch_margin_tiles(self, dem, horizons_deg, body_vectors_m, *, tile_y, tile_x,
                                valid_height=PATCH_SIZE, valid_width=PATCH_SIZE
                                ) -> Iterator[np.ndarray]:
        for batch in self._signal_batches(dem, horizons_deg, body_vectors_m, tile_y=tile_y,
                                          tile_x=tile_x, valid_height=valid_height,
                                          valid_width=valid_width, margin=True):
            for tile in batch:
                yield np.ascontiguousarray(tile)

    def iter_patch_tiles(self, dem, horizons_deg, sun_vectors_m, *, tile_y, tile_x,
                         valid_height=PATCH_SIZE, valid_width=PATCH_SIZE) -> Iterator[np.ndarray]:
        for batch in self._signal_batches(dem, horizons_deg, sun_vectors_m, tile_y=tile_y,
                                          tile_x=tile_x, valid_height=valid_height,
                                          valid_width=valid_width, margin=False):
            encoded = (batch * np.float32(255.0)).astype(np.uint8)
            for tile in encoded:
                yield np.ascontiguousarray(tile)

    def compute_patch(self, dem, horizons_deg, reduced_sun_vectors_m, *, tile_y, tile_x,
                      valid_height=PATCH_SIZE, valid_width=PATCH_SIZE) -> np.ndarray:
        """uint8 PSR mask [valid_height, valid_width] (255 = PSR)."""
        self._check_patch(dem, tile_y, tile_x, valid_height, valid_width)
        vectors = validate_vectors(reduced_sun_vectors_m)
        hz = self._upload_patch(horizons_deg)
        out = self.lighting_tiles(dem, [(tile_x, tile_y)], hz, psr_vectors=vectors,
                                  sun_mean=False, earth_visible=False, psr=True)
        try:
            mask = out.psr.to_host().reshape(PATCH_SIZE, PATCH_SIZE)
        finally:
            out.free()
        return np.ascontiguousarray(mask[:valid_height, :valid_width])


class CpuLightingSession:
    """The same lighting on host threads: the ``"cpu"`` backend.

    Drop-in for the reference's ``LightmapCpuSession`` (lightmap_cpu.py:193-237)
    and its CPU PSR path (``compute_psr_patch_reference``, psr.py:94-140). The
    arithmetic is the HIP kernel's (``lsh_lighting_core.h``) compiled for the
    host, so the two backends agree up to the host and device math libraries.
    It never touches a GPU.
    """

    def __init__(self, *, threads: int = 0, time_batch_size: int = 64,
                 sun_half_angle_deg: float = SUN_HALF_ANGLE_DEG,
                 sun_angular_size_deg: float = SUN_ANGULAR_SIZE_DEG):
        if time_batch_size < 1:
            raise ValueError("time_batch_size must be positive")
        self.lib = _lib.load_library()
        self.threads = int(threads)
        self.time_batch_size = int(time_batch_size)
        self.sun_half_angle_deg = float(sun_half_angle_deg)
        self.sun_angular_size_deg = float(sun_angular_size_deg)

    def lighting_tiles(self, dem: DemGrid, tiles, horizons: np.ndarray, *, sun=None, earth=None,
                       psr_vectors=None, azimuth_major: bool = False, sun_bytes: bool = False,
                       sun_fraction: bool = False, sun_mean: bool = True,
                       earth_margin: bool = False, earth_visible: bool = True, psr: bool = True,
                       psr_margin: bool = False) -> dict[str, np.ndarray]:
        """Host arrays in, host arrays out; shapes as ``HipLightingSession.lighting_tiles``."""
        tile_arr = _tile_array(tiles)
        n = len(tile_arr)
        hz = np.ascontiguousarray(horizons, dtype=np.float32)
        N = hz.shape[-2] if azimuth_major else hz.shape[-1]
        if hz.size != n * PATCH_PIXELS * N:
            raise ValueError("horizons must hold [tile][16384][N] (or [tile][N][16384]) values")
        vecs = {k: (validate_vectors(v) if v is not None else None)
                for k, v in (("sun", sun), ("earth", earth), ("psr", psr_vectors))}
        T = 0 if vecs["sun"] is None else len(vecs["sun"])
        E = 0 if vecs["earth"] is None else len(vecs["earth"])
        P = 0 if vecs["psr"] is None else len(vecs["psr"])
        want = {
            "sun_bytes": (sun_bytes and T, (n, T, PATCH_PIXELS), np.uint8, _lib.LIGHT_SUN_BYTES),
            "sun_fraction": (sun_fraction and T, (n, T, PATCH_PIXELS), np.float32, _lib.LIGHT_SUN_FRACTION),
            "sun_mean": (sun_mean and T, (n, PATCH_PIXELS), np.float32, _lib.LIGHT_SUN_MEAN),
            "earth_margin": (earth_margin and E, (n, E, PATCH_PIXELS), np.float32, _lib.LIGHT_EARTH_MARGIN),
            "earth_visible": (earth_visible and E, (n, PATCH_PIXELS), np.float32, _lib.LIGHT_EARTH_VISIBLE),
            "psr": (psr and P, (n, PATCH_PIXELS), np.uint8, _lib.LIGHT_PSR),
            "psr_margin": (psr_margin and P, (n, PATCH_PIXELS), np.float32, _lib.LIGHT_PSR_MARGIN),
        }
        out, ptrs, flags = {}, {}, 0
        for name, (enabled, shape, dtype, flag) in want.items():
            if enabled:
                out[name] = np.empty(shape, dtype)
                ptrs[name] = out[name].ctypes.data
                flags |= flag
            else:
                ptrs[name] = None
        elevation = np.ascontiguousarray(dem.elevation_m, dtype=np.float32)
        gt = np.ascontiguousarray(dem.geo_transform, dtype=np.float64)
        proj = projection_array(dem.projection)

        def ptr(v):
            return None if v is None else v.ctypes.data

        check(self.lib.lsh_lighting_tiles_host(
            elevation.ctypes.data, dem.width, dem.height, gt.ctypes.data, proj.ctypes.data,
            tile_arr.ctypes.data, n, hz.ctypes.data, N, 1 if azimuth_major else 0,
            ptr(vecs["sun"]), T, ptr(vecs["earth"]), E, ptr(vecs["psr"]), P,
            self.sun_half_angle_deg, self.sun_angular_size_deg, flags, self.threads,
            ptrs["sun_bytes"], ptrs["sun_fraction"], ptrs["sun_mean"], ptrs["earth_margin"],
            ptrs["earth_visible"], ptrs["psr"], ptrs["psr_margin"]), "lsh_lighting_tiles_host")
        return out

    def _signal_batches(self, dem, horizons_deg, vectors_m, *, tile_y, tile_x, valid_height,
                        valid_width, margin: bool):
        HipLightingSession._check_patch(dem, tile_y, tile_x, valid_height, valid_width)
        vectors = validate_vectors(vectors_m)
        hz = _finite_patch(horizons_deg)
        for start in range(0, len(vectors), self.time_batch_size):
            batch = vectors[start:start + self.time_batch_size]
            if margin:
                values = self.lighting_tiles(dem, [(tile_x, tile_y)], hz, earth=batch,
                                             sun_mean=False, earth_visible=False, psr=False,
                                             earth_margin=True)["earth_margin"]
            else:
                values = self.lighting_tiles(dem, [(tile_x, tile_y)], hz, sun=batch,
                                             sun_mean=False, earth_visible=False, psr=False,
                                             sun_fraction=True)["sun_fraction"]
            yield values.reshape(len(batch), PATCH_SIZE, PATCH_SIZE)[:, :valid_height, :valid_width]

    def duration_tiles(self, dem: DemGrid, tiles, horizons: np.ndarray, plan, sun, earth=None, *,
                       azimuth_major: bool = False) -> np.ndarray:
        """``HipLightingSession.duration_tiles`` on host threads: float32 [bands][tile][16384]."""
        tile_arr = _tile_array(tiles)
        n = len(tile_arr)
        hz = np.ascontiguousarray(horizons, dtype=np.float32)
        N = hz.shape[-2] if azimuth_major else hz.shape[-1]
        if hz.size != n * PATCH_PIXELS * N:
            raise ValueError("horizons must hold [tile][16384][N] (or [tile][N][16384]) values")
        sun_v = validate_vectors(sun)
        earth_v = validate_vectors(earth) if (earth is not None and plan.needs_earth) else None
        spec = plan.spec()
        need = ctypes.c_int64()
        check(self.lib.lsh_duration_workspace_bytes(ctypes.byref(spec), n, ctypes.byref(need)),
              "lsh_duration_workspace_bytes")
        workspace = np.empty(max(int(need.value), 8), np.uint8)
        out = np.empty((plan.bands, n, PATCH_PIXELS), np.float32)
        elevation = np.ascontiguousarray(dem.elevation_m, dtype=np.float32)
        gt = np.ascontiguousarray(dem.geo_transform, dtype=np.float64)
        proj = projection_array(dem.projection)
        check(self.lib.lsh_duration_tiles_host(
            elevation.ctypes.data, dem.width, dem.height, gt.ctypes.data, proj.ctypes.data,
            tile_arr.ctypes.data, n, hz.ctypes.data, N, 1 if azimuth_major else 0,
            sun_v.ctypes.data, None if earth_v is None else earth_v.ctypes.data,
            self.sun_half_angle_deg, ctypes.byref(spec), self.threads, workspace.ctypes.data,
            out.ctypes.data), "lsh_duration_tiles_host")
        return out

    def duration_patch(self, dem, horizons_deg, plan, sun_vectors_m, earth_vectors_m=None, *,
                       tile_y, tile_x, valid_height=PATCH_SIZE, valid_width=PATCH_SIZE):
        """One patch's bands, float32 [valid_height, valid_width] each."""
        HipLightingSession._check_patch(dem, tile_y, tile_x, valid_height, valid_width)
        out = self.duration_tiles(dem, [(tile_x, tile_y)], _finite_patch(horizons_deg), plan,
                                  sun_vectors_m, earth_vectors_m)
        values = out.reshape(plan.bands, PATCH_SIZE, PATCH_SIZE)
        return tuple(np.ascontiguousarray(b[:valid_height, :valid_width]) for b in values)

    iter_patch_fraction_tiles = HipLightingSession.iter_patch_fraction_tiles
    iter_patch_margin_tiles = HipLightingSession.iter_patch_margin_tiles
    iter_patch_tiles = HipLightingSession.iter_patch_tiles

    def compute_patch(self, dem, horizons_deg, reduced_sun_vectors_m, *, tile_y, tile_x,
                      valid_height=PATCH_SIZE, valid_width=PATCH_SIZE) -> np.ndarray:
        """uint8 PSR mask [valid_height, valid_width] (255 = PSR)."""
        HipLightingSession._check_patch(dem, tile_y, tile_x, valid_height, valid_width)
        vectors = validate_vectors(reduced_sun_vectors_m)
        mask = self.lighting_tiles(dem, [(tile_x, tile_y)], _finite_patch(horizons_deg),
                                   psr_vectors=vectors, sun_mean=False, earth_visible=False,
                                   psr=True)["psr"].reshape(PATCH_SIZE, PATCH_SIZE)
        return np.ascontiguousarray(mask[:valid_height, :valid_width])


def _finite_patch(horizons_deg) -> np.ndarray:
    horizons = np.ascontiguousarray(horizons_deg, dtype=np.float32)
    if horizons.ndim != 3 or horizons.shape[:2] != (PATCH_SIZE, PATCH_SIZE):
        raise ValueError("horizons must have shape (128, 128, N)")
    if not np.all(np.isfinite(horizons)):
        raise ValueError("horizons must contain finite elevations")
    return horizons.reshape(1, PATCH_PIXELS, horizons.shape[2])


def select_lighting_session(backend: str = "auto", **options):
    """The lighting calculator for a backend (lightmap_pipeline.py:81-108, psr_pipeline.py:218-258).

    ``"cpu"`` never probes the GPU; ``"hip"`` never falls back (a missing
    library or device raises ``HipBackendError``); ``"auto"`` uses the GPU when a
    session can be created and the host path otherwise. Returns
    ``(session, selected_backend)``.
    """
    from .errors import HipBackendError

    if backend not in ("auto", "cpu", "hip"):
        raise ValueError("backend must be 'auto', 'cpu', or 'hip'")
    host_options = {k: v for k, v in options.items() if k != "device_id"}
    if backend == "cpu":
        return CpuLightingSession(**host_options), "cpu"
    try:
        return HipLightingSession(**options), "hip"
    except HipBackendError:
        if backend == "hip":
            raise
        return CpuLightingSession(**host_options), "cpu"
