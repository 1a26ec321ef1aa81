"""Horizon tile files: the naming contract in Python, the file rules in C++.

A tile holds 128 x 128 pixels x 1440 azimuths of float32 horizon degrees,
pixel-major. The contract shared with the reference's C#/Python tile store
(file_format.py:18-30, 283-415), so that either side reads the other's tiles:

* name ``horizon_{y:05d}_{x:05d}_{dm:03d}.{cbin|bin}`` in the partition
  directory ``{y:05d}``, with dm the observer height in decimetres truncated
  toward zero (0..999) and tile coordinates 0..99999;
* a tile is looked up as partitioned .cbin, partitioned .bin and then, unless
  disabled, the legacy flat .cbin and .bin in the root;
* only a structurally complete file counts as present (the resume rule);
* pixels outside a partial edge tile hold -50 degrees, the .cbin minimum.

What a file must contain to be complete, how it is read and how it is
published (staged beside the destination, fsynced, renamed) is C++ in the
host library beside the codec (``lsh_tile_file_*``, ``lsh_cbin_*``); this
module resolves names and pads edge tiles.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _lib

PATCH_SIZE = 128
AZIMUTH_COUNT = 1440
PIXEL_COUNT = PATCH_SIZE * PATCH_SIZE
RAW_FILE_BYTES = PIXEL_COUNT * AZIMUTH_COUNT * 4
EDGE_PADDING_DEG = np.float32(-50.0)
TILE_SHAPE = (PATCH_SIZE, PATCH_SIZE, AZIMUTH_COUNT)


@dataclass(frozen=True)
class HorizonReadTimings:
    """Seconds spent reading the file and decoding it (file_format.py:31-36)."""

    file_read_seconds: float
    decompression_seconds: float


def _raise_for(status: int, what: str) -> None:
    if status == _lib.LSH_EINVAL:
        raise ValueError(f"{what}: {_lib.last_error()}")
    _lib.check(status, what)


# -- codec ------------------------------------------------------------------------------------
def encode_cbin(degrees: np.ndarray, *, threads: int = 0) -> bytes:
    """``.cbin`` payload of float32 degrees ``[pixels, azimuths]``."""
    values = np.ascontiguousarray(degrees, dtype=np.float32)
    if values.ndim != 2:
        raise ValueError("degrees must be [pixels, azimuths]")
    pixels, azimuths = values.shape
    buffer = np.empty(pixels * (2 * azimuths + 2), dtype=np.uint8)
    used = ctypes.c_int64()
    _raise_for(_lib.load_library().lsh_cbin_encode(values.ctypes.data, pixels, azimuths,
                                                   buffer.ctypes.data, buffer.size,
                                                   ctypes.byref(used), threads),
               "lsh_cbin_encode")
    return buffer[: used.value].tobytes()


def decode_cbin(payload, *, pixels: int = PIXEL_COUNT, azimuths: int = AZIMUTH_COUNT,
                out: np.ndarray | None = None, threads: int = 0) -> np.ndarray:
    """float32 ``[pixels, azimuths]`` of a complete payload; ValueError names the first fault."""
    data = np.ascontiguousarray(np.frombuffer(payload, dtype=np.uint8)
                                if not isinstance(payload, np.ndarray) else payload, dtype=np.uint8)
    target = _writable_float_tile(out, pixels * azimuths) if out is not None else \
        np.empty(pixels * azimuths, dtype=np.float32)
    _raise_for(_lib.load_library().lsh_cbin_decode(data.ctypes.data, data.size, pixels, azimuths,
                                                   target.ctypes.data, threads), "lsh_cbin_decode")
    return target.reshape(pixels, azimuths)


def cbin_is_complete(payload, *, pixels: int = PIXEL_COUNT, azimuths: int = AZIMUTH_COUNT) -> bool:
    data = np.ascontiguousarray(np.frombuffer(payload, dtype=np.uint8))
    flag = ctypes.c_int32()
    _raise_for(_lib.load_library().lsh_cbin_complete(data.ctypes.data, data.size, pixels,
                                                     azimuths, ctypes.byref(flag)),
               "lsh_cbin_complete")
    return flag.value == 1


def _writable_float_tile(out: np.ndarray, count: int) -> np.ndarray:
    if (not isinstance(out, np.ndarray) or out.dtype != np.float32 or out.size != count
            or not out.flags.c_contiguous or not out.flags.writeable):
        raise ValueError("output must be a writable contiguous float32 array of the tile's size")
    return out.reshape(-1)


# -- edge tiles -------------------------------------------------------------------------------
def pad_tile(degrees, *, valid_width: int = PATCH_SIZE, valid_height: int = PATCH_SIZE) -> np.ndarray:
    """The full float32 ``[16384, 1440]`` tile of a valid region (file_format.py:93-120).

    ``degrees`` is either the valid region ``[valid_height * valid_width, 1440]``
    or a full tile whose pixels outside the region are replaced; those pixels
    hold -50 degrees. NaN is rejected.
    """
    if not (1 <= valid_width <= PATCH_SIZE and 1 <= valid_height <= PATCH_SIZE):
        raise ValueError("valid tile dimensions must be between 1 and 128")
    values = np.asarray(degrees, dtype=np.float32)
    region_shape = (valid_height * valid_width, AZIMUTH_COUNT)
    if values.shape == (PIXEL_COUNT, AZIMUTH_COUNT):
        region = values.reshape(TILE_SHAPE)[:valid_height, :valid_width]
    elif values.shape == region_shape:
        region = values.reshape(valid_height, valid_width, AZIMUTH_COUNT)
    else:
        raise ValueError(f"degrees must have shape {region_shape} or {(PIXEL_COUNT, AZIMUTH_COUNT)}")
    if np.isnan(region).any():
        raise ValueError("degrees must not contain NaN values")
    if region.shape[:2] == (PATCH_SIZE, PATCH_SIZE):
        return np.ascontiguousarray(region.reshape(PIXEL_COUNT, AZIMUTH_COUNT))
    full = np.pad(region, ((0, PATCH_SIZE - valid_height), (0, PATCH_SIZE - valid_width), (0, 0)),
                  constant_values=EDGE_PADDING_DEG)
    return np.ascontiguousarray(full.reshape(PIXEL_COUNT, AZIMUTH_COUNT))


normalize_tile = pad_tile  # the reference's name for the same operation


# -- files ------------------------------------------------------------------------------------
def tile_file_status(path) -> int:
    """``_lib.TILE_ABSENT``, ``TILE_COMPLETE`` or ``TILE_INCOMPLETE`` for one file."""
    status = ctypes.c_int32()
    _raise_for(_lib.load_library().lsh_tile_file_status(str(path).encode(), PIXEL_COUNT,
                                                        AZIMUTH_COUNT, ctypes.byref(status)),
               "lsh_tile_file_status")
    return status.value


def read_horizon_tile_with_timings(path, *, output: np.ndarray | None = None):
    """float32 ``[128, 128, 1440]`` of one tile file and its read / decode seconds."""
    if output is not None:
        if output.shape != TILE_SHAPE:
            raise ValueError("output must be a writable contiguous float32 (128, 128, 1440) array")
        target = _writable_float_tile(output, PIXEL_COUNT * AZIMUTH_COUNT)
    else:
        target = np.empty(PIXEL_COUNT * AZIMUTH_COUNT, dtype=np.float32)
    seconds = np.zeros(2, dtype=np.float64)
    _raise_for(_lib.load_library().lsh_tile_file_read(str(path).encode(), PIXEL_COUNT,
                                                      AZIMUTH_COUNT, target.ctypes.data, 0,
                                                      seconds.ctypes.data),
               "horizon tile read")
    return target.reshape(TILE_SHAPE), HorizonReadTimings(float(seconds[0]), float(seconds[1]))


def read_horizon_tile(path) -> np.ndarray:
    """float32 ``[128, 128, 1440]`` of one tile file (file_format.py:277-280)."""
    return read_horizon_tile_with_timings(path)[0]


def write_tile_file(path, tile: np.ndarray, *, compress: bool) -> Path:
    """Publish a full float32 ``[16384, 1440]`` tile at ``path`` (staged, fsynced, renamed)."""
    values = np.ascontiguousarray(tile, dtype=np.float32)
    if values.shape != (PIXEL_COUNT, AZIMUTH_COUNT):
        raise ValueError("a tile file holds [16384, 1440] float32 values")
    target = Path(path)
    target.parent.mkdir(parents=True, exist_ok=True)
    _raise_for(_lib.load_library().lsh_tile_file_write(str(target).encode(), values.ctypes.data,
                                                       PIXEL_COUNT, AZIMUTH_COUNT,
                                                       1 if compress else 0, 0),
               "horizon tile write")
    return target


# -- names and the store ----------------------------------------------------------------------
def _decimetres(observer_elevation_m: float) -> int:
    dm = int(float(observer_elevation_m) * 10.0)  # truncation toward zero, as the C# cast
    if not 0 <= dm <= 999:
        raise ValueError("observer elevation must fit the D3 filename field")
    return dm


def _tile_coordinate(value) -> int:
    if isinstance(value, bool) or not isinstance(value, (int, np.integer)) or not 0 <= value <= 99999:
        raise ValueError("horizon tile coordinates must fit the D5 filename field")
    return int(value)


def tile_file_name(tile_y: int, tile_x: int, observer_elevation_m: float, *, compress: bool) -> str:
    y, x, dm = _tile_coordinate(tile_y), _tile_coordinate(tile_x), _decimetres(observer_elevation_m)
    return "horizon_%05d_%05d_%03d.%s" % (y, x, dm, "cbin" if compress else "bin")


class HorizonTileStore:
    """A horizon tile directory with the reference store's interface (file_format.py:283-415)."""

    def __init__(self, root_directory, *, read_legacy_flat_files: bool = True):
        if not str(root_directory):
            raise ValueError("horizon root directory must be provided")
        self.root = Path(root_directory)
        self.read_legacy_flat_files = bool(read_legacy_flat_files)

    def build_file_name(self, tile_y: int, tile_x: int, observer_elevation_m: float, *,
                        compress: bool) -> str:
        return tile_file_name(tile_y, tile_x, observer_elevation_m, compress=compress)

    def build_path(self, tile_y: int, tile_x: int, observer_elevation_m: float, *,
                   compress: bool) -> Path:
        name = tile_file_name(tile_y, tile_x, observer_elevation_m, compress=compress)
        return self.root / ("%05d" % int(tile_y)) / name

    def candidate_paths(self, tile_y: int, tile_x: int,
                        observer_elevation_m: float) -> tuple[Path, ...]:
        names = [tile_file_name(tile_y, tile_x, observer_elevation_m, compress=c)
                 for c in (True, False)]
        directories = [self.root / ("%05d" % int(tile_y))]
        if self.read_legacy_flat_files:
            directories.append(self.root)
        return tuple(d / n for d in directories for n in names)

    @staticmethod
    def is_complete(path) -> bool:
        return tile_file_status(path) == _lib.TILE_COMPLETE

    def find_existing_path(self, tile_y: int, tile_x: int, observer_elevation_m: float, *,
                           require_complete: bool = True) -> Path | None:
        wanted = {_lib.TILE_COMPLETE} if require_complete else {_lib.TILE_COMPLETE,
                                                                _lib.TILE_INCOMPLETE}
        return next((p for p in self.candidate_paths(tile_y, tile_x, observer_elevation_m)
                     if tile_file_status(p) in wanted), None)

    def read(self, tile_y: int, tile_x: int, observer_elevation_m: float) -> np.ndarray | None:
        """The preferred complete tile as float32 [128, 128, 1440], or None."""
        path = self.find_existing_path(tile_y, tile_x, observer_elevation_m)
        return None if path is None else read_horizon_tile(path)

    def write(self, tile_y: int, tile_x: int, observer_elevation_m: float, degrees, *,
              compress: bool, valid_width: int = PATCH_SIZE, valid_height: int = PATCH_SIZE) -> Path:
        path = self.build_path(tile_y, tile_x, observer_elevation_m, compress=compress)
        tile = pad_tile(degrees, valid_width=valid_width, valid_height=valid_height)
        return write_tile_file(path, tile, compress=compress)


__all__ = [
    "AZIMUTH_COUNT", "PATCH_SIZE", "PIXEL_COUNT", "RAW_FILE_BYTES", "HorizonReadTimings",
    "HorizonTileStore", "cbin_is_complete", "decode_cbin", "encode_cbin", "normalize_tile",
    "pad_tile", "read_horizon_tile", "read_horizon_tile_with_timings", "tile_file_name",
    "tile_file_status", "write_tile_file",
]
