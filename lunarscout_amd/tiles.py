"""Horizon tile files: naming, completeness, the ``.cbin`` codec and staged writes.

Mirrors the reference's production tile store (``file_format.py:283-415``) so
tiles written here are read by the reference and the other way round:

* path ``root/{y:05d}/horizon_{y:05d}_{x:05d}_{dm:03d}.{cbin|bin}`` with
  ``dm = trunc(observer height * 10)`` (``:293-321``);
* lookup order partitioned ``.cbin``, partitioned ``.bin``, then the legacy
  flat files (``:323-333``);
* resume rule "structurally complete" (``:335-356``): a ``.bin`` of exactly
  the raw size (or raw + 28 legacy metadata bytes), a ``.cbin`` whose blocks
  all parse with nothing after the last;
* writes staged beside the destination, fsynced and atomically renamed
  (``:378-415``); edge tiles padded to 128x128 with -50 degrees (``:93-120``).

The codec itself is the C++ host code of ``liblunarscout_hip.so``
(``lsh_cbin_*``), multi-threaded; it touches no device.
"""

from __future__ import annotations

import ctypes
import os
import time
from dataclasses import dataclass
from pathlib import Path
from uuid import uuid4

import numpy as np

from . import _lib

PATCH_SIZE = 128
AZIMUTH_COUNT = 1440
PIXEL_COUNT = PATCH_SIZE * PATCH_SIZE
TOTAL_SAMPLES = PIXEL_COUNT * AZIMUTH_COUNT
RAW_FILE_BYTES = TOTAL_SAMPLES * 4
LEGACY_RAW_FILE_BYTES = RAW_FILE_BYTES + 28
MAX_COMPRESSED_BLOCK_BYTES = 2 * AZIMUTH_COUNT
EDGE_PADDING_DEG = -50.0


@dataclass(frozen=True)
class HorizonReadTimings:
    """Wall-clock split of one tile read (``file_format.py:31-36``)."""

    file_read_seconds: float
    decompression_seconds: float


def encode_cbin(degrees: np.ndarray, *, threads: int = 0) -> bytes:
    """``.cbin`` payload of float32 degrees ``[pixels, azimuths]``."""
    values = np.ascontiguousarray(degrees, dtype=np.float32)
    if values.ndim != 2:
        raise ValueError("degrees must be [pixels, azimuths]")
    pixels, azimuths = values.shape
    capacity = pixels * (2 * azimuths + 2)
    out = np.empty(capacity, dtype=np.uint8)
    size = ctypes.c_int64()
    status = _lib.load_library().lsh_cbin_encode(values.ctypes.data, pixels, azimuths,
                                                 out.ctypes.data, capacity, ctypes.byref(size),
                                                 threads)
    _codec_check(status, "lsh_cbin_encode")
    return out[: size.value].tobytes()


def decode_cbin(payload, *, pixels: int = PIXEL_COUNT, azimuths: int = AZIMUTH_COUNT,
                out: np.ndarray | None = None, threads: int = 0) -> np.ndarray:
    """float32 ``[pixels, azimuths]`` from a complete ``.cbin`` payload.

    Raises ``ValueError`` naming the first structural error (``file_format.py:135-182``).
    """
    data = np.frombuffer(payload, dtype=np.uint8) if not isinstance(payload, np.ndarray) else payload
    data = np.ascontiguousarray(data, dtype=np.uint8)
    if out is None:
        out = np.empty((pixels, azimuths), dtype=np.float32)
    elif out.dtype != np.float32 or out.size != pixels * azimuths or not out.flags.c_contiguous:
        raise ValueError("output must be a contiguous float32 array of the tile's size")
    status = _lib.load_library().lsh_cbin_decode(data.ctypes.data, data.size, pixels, azimuths,
                                                 out.ctypes.data, threads)
    _codec_check(status, "lsh_cbin_decode")
    return out.reshape(pixels, azimuths)


def cbin_is_complete(payload, *, pixels: int = PIXEL_COUNT, azimuths: int = AZIMUTH_COUNT) -> bool:
    data = np.ascontiguousarray(np.frombuffer(payload, dtype=np.uint8))
    flag = ctypes.c_int32()
    status = _lib.load_library().lsh_cbin_complete(data.ctypes.data, data.size, pixels, azimuths,
                                                   ctypes.byref(flag))
    _codec_check(status, "lsh_cbin_complete")
    return bool(flag.value)


def _codec_check(status: int, what: str) -> None:
    if status == _lib.LSH_EINVAL:
        raise ValueError(f"{what}: {_lib.last_error()}")
    _lib.check(status, what)


def normalize_tile(degrees, *, valid_width: int = PATCH_SIZE,
                   valid_height: int = PATCH_SIZE) -> np.ndarray:
    """A complete float32 ``[16384, 1440]`` tile (``_normalize_tile``, file_format.py:93-120).

    Accepts the valid region ``[vh * vw, 1440]`` or a full tile; pixels outside
    the valid region become -50 degrees; NaN is rejected.
    """
    if not 1 <= valid_width <= PATCH_SIZE or not 1 <= valid_height <= PATCH_SIZE:
        raise ValueError("valid tile dimensions must be between 1 and 128")
    source = np.asarray(degrees, dtype=np.float32)
    expected = (valid_width * valid_height, AZIMUTH_COUNT)
    full = (PIXEL_COUNT, AZIMUTH_COUNT)
    if source.shape not in (expected, full):
        raise ValueError(f"degrees must have shape {expected} or {full}")
    if np.isnan(source).any():
        raise ValueError("degrees must not contain NaN values")
    if valid_width == PATCH_SIZE and valid_height == PATCH_SIZE:
        return np.ascontiguousarray(source)
    padded = np.full((PATCH_SIZE, PATCH_SIZE, AZIMUTH_COUNT), EDGE_PADDING_DEG, dtype=np.float32)
    if source.shape == full:
        region = source.reshape(PATCH_SIZE, PATCH_SIZE, AZIMUTH_COUNT)[:valid_height, :valid_width]
    else:
        region = source.reshape(valid_height, valid_width, AZIMUTH_COUNT)
    padded[:valid_height, :valid_width] = region
    return padded.reshape(full)


def read_horizon_tile_with_timings(path, *, output: np.ndarray | None = None):
    """float32 ``[128, 128, 1440]`` of one tile and its read / decode times (file_format.py:219-274)."""
    flat = None
    if output is not None:
        if (output.dtype != np.float32 or output.shape != (PATCH_SIZE, PATCH_SIZE, AZIMUTH_COUNT)
                or not output.flags.c_contiguous or not output.flags.writeable):
            raise ValueError("output must be a writable contiguous float32 (128, 128, 1440) array")
        flat = output.reshape(PIXEL_COUNT, AZIMUTH_COUNT)
    candidate = Path(path)
    suffix = candidate.suffix.lower()
    started = time.perf_counter()
    if suffix == ".bin":
        try:
            size = candidate.stat().st_size
            if size not in (RAW_FILE_BYTES, LEGACY_RAW_FILE_BYTES):
                raise ValueError("uncompressed horizon file size is invalid")
            with candidate.open("rb") as handle:
                payload = handle.read(RAW_FILE_BYTES)
        except OSError as exc:
            raise ValueError(f"unable to read horizon tile: {candidate}") from exc
        read_seconds = time.perf_counter() - started
        decode_started = time.perf_counter()
        source = np.frombuffer(payload, dtype="<f4", count=TOTAL_SAMPLES)
        if flat is None:
            values = source.astype(np.float32).reshape(PIXEL_COUNT, AZIMUTH_COUNT)
        else:
            np.copyto(flat.reshape(-1), source)
            values = flat
    elif suffix == ".cbin":
        try:
            payload = candidate.read_bytes()
        except OSError as exc:
            raise ValueError(f"unable to read horizon tile: {candidate}") from exc
        read_seconds = time.perf_counter() - started
        decode_started = time.perf_counter()
        values = decode_cbin(payload, out=flat)
    else:
        raise ValueError("horizon tile must have a .bin or .cbin extension")
    decode_seconds = time.perf_counter() - decode_started
    return (values.reshape(PATCH_SIZE, PATCH_SIZE, AZIMUTH_COUNT),
            HorizonReadTimings(read_seconds, decode_seconds))


def read_horizon_tile(path) -> np.ndarray:
    """float32 ``[128, 128, 1440]`` of one complete tile (file_format.py:277-280)."""
    return read_horizon_tile_with_timings(path)[0]


class HorizonTileStore:
    """The production tile directory (``HorizonTileStore``, file_format.py:283-415)."""

    def __init__(self, root_directory, *, read_legacy_flat_files: bool = True):
        self.root = Path(root_directory)
        if not str(root_directory):
            raise ValueError("horizon root directory must be provided")
        self.read_legacy_flat_files = bool(read_legacy_flat_files)

    @staticmethod
    def _elevation_decimeters(observer_elevation_m: float) -> int:
        value = int(float(observer_elevation_m) * 10.0)  # truncation toward zero
        if not 0 <= value <= 999:
            raise ValueError("observer elevation must fit the D3 filename field")
        return value

    @staticmethod
    def _coordinate(value: int) -> int:
        if isinstance(value, bool) or not isinstance(value, (int, np.integer)) or not 0 <= value <= 99999:
            raise ValueError("horizon tile coordinates must fit the D5 filename field")
        return int(value)

    def build_file_name(self, tile_y: int, tile_x: int, observer_elevation_m: float, *,
                        compress: bool) -> str:
        y, x = self._coordinate(tile_y), self._coordinate(tile_x)
        dm = self._elevation_decimeters(observer_elevation_m)
        return f"horizon_{y:05d}_{x:05d}_{dm:03d}{'.cbin' if compress else '.bin'}"

    def build_path(self, tile_y: int, tile_x: int, observer_elevation_m: float, *,
                   compress: bool) -> Path:
        name = self.build_file_name(tile_y, tile_x, observer_elevation_m, compress=compress)
        return self.root / f"{int(tile_y):05d}" / name

    def candidate_paths(self, tile_y: int, tile_x: int, observer_elevation_m: float) -> tuple[Path, ...]:
        names = [self.build_file_name(tile_y, tile_x, observer_elevation_m, compress=c)
                 for c in (True, False)]
        partition = self.root / f"{int(tile_y):05d}"
        paths = [partition / n for n in names]
        if self.read_legacy_flat_files:
            paths += [self.root / n for n in names]
        return tuple(paths)

    @staticmethod
    def is_complete(path) -> bool:
        candidate = Path(path)
        try:
            suffix = candidate.suffix.lower()
            if suffix == ".bin":
                return candidate.stat().st_size in (RAW_FILE_BYTES, LEGACY_RAW_FILE_BYTES)
            if suffix != ".cbin" or candidate.stat().st_size <= 0:
                return False
            return cbin_is_complete(candidate.read_bytes())
        except (OSError, ValueError):
            return False

    def find_existing_path(self, tile_y: int, tile_x: int, observer_elevation_m: float, *,
                           require_complete: bool = True) -> Path | None:
        for path in self.candidate_paths(tile_y, tile_x, observer_elevation_m):
            if path.is_file() and (not require_complete or self.is_complete(path)):
                return path
        return None

    def read(self, tile_y: int, tile_x: int, observer_elevation_m: float) -> np.ndarray | None:
        path = self.find_existing_path(tile_y, tile_x, observer_elevation_m)
        return None if path is None else read_horizon_tile(path)

    def write(self, tile_y: int, tile_x: int, observer_elevation_m: float, degrees, *,
              compress: bool, valid_width: int = PATCH_SIZE, valid_height: int = PATCH_SIZE) -> Path:
        final = self.build_path(tile_y, tile_x, observer_elevation_m, compress=compress)
        final.parent.mkdir(parents=True, exist_ok=True)
        extension = ".cbin" if compress else ".bin"
        staging = final.parent / f"{final.stem}.{uuid4().hex}.tmp{extension}"
        tile = normalize_tile(degrees, valid_width=valid_width, valid_height=valid_height)
        payload = encode_cbin(tile) if compress else tile.astype("<f4", copy=False).tobytes()
        try:
            with staging.open("wb", buffering=1 << 20) as handle:
                handle.write(payload)
                handle.flush()
                os.fsync(handle.fileno())
            os.replace(staging, final)
        except BaseException:
            try:
                staging.unlink(missing_ok=True)
            except OSError:
                pass
            raise
        return final


__all__ = [
    "AZIMUTH_COUNT", "PATCH_SIZE", "PIXEL_COUNT", "RAW_FILE_BYTES", "HorizonReadTimings",
    "HorizonTileStore", "cbin_is_complete", "decode_cbin", "encode_cbin", "normalize_tile",
    "read_horizon_tile", "read_horizon_tile_with_timings",
]
