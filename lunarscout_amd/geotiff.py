"""Minimal GeoTIFF reader and writer for stereographic lunar DEMs and products.

The reference reads and writes rasters through rasterio/GDAL and interprets the
CRS with pyproj (``src/lunarscout/geotiff.py:60-151``, ``products.py:97-177``);
neither is available here, so this module reads the TIFF structure itself:

* classic TIFF (``II*`` / ``MM*``), one band, strips or tiles;
* compression none (1) or deflate (8, 32946); predictor none (1), horizontal
  (2) or floating point (3, byte planes most significant first);
* sample formats unsigned / signed integer and IEEE float, 8-64 bits;
* georeferencing from ModelTiepoint + ModelPixelScale or ModelTransformation
  (half-pixel shift for PixelIsPoint, as GDAL applies it) and the
  stereographic GeoKeys GDAL writes (coordinate transformations 14, 15, 16:
  origin latitude / longitude, scale, false easting / northing, semi-major
  axis);
* GDAL_NODATA.

``write_geotiff`` writes what the reference writes for its products
(``geotiff.py:190-199``): tiled 128x128, deflate, predictor 3 for floats and 2
for integers, with the same stereographic GeoKeys.
"""

from __future__ import annotations

import math
import os
import struct
import tempfile
import zlib
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .errors import InputError, LunarscoutError
from .geometry import DemGrid, ProjectionParameters

# TIFF tags
_WIDTH, _HEIGHT, _BITS, _COMPRESSION, _PHOTOMETRIC = 256, 257, 258, 259, 262
_STRIP_OFFSETS, _SAMPLES, _ROWS_PER_STRIP, _STRIP_COUNTS = 273, 277, 278, 279
_PLANAR, _PREDICTOR = 284, 317
_TILE_W, _TILE_H, _TILE_OFFSETS, _TILE_COUNTS = 322, 323, 324, 325
_SAMPLE_FORMAT = 339
_PIXEL_SCALE, _TIEPOINT, _TRANSFORMATION = 33550, 33922, 34264
_GEOKEYS, _GEO_DOUBLES, _GEO_ASCII, _GDAL_NODATA = 34735, 34736, 34737, 42113
_GDAL_METADATA = 42112

# GeoKeys
_K_MODEL_TYPE, _K_RASTER_TYPE = 1024, 1025
_K_GEOG_TYPE, _K_DATUM, _K_ANGULAR, _K_ELLIPSOID = 2048, 2050, 2054, 2056
_K_SEMI_MAJOR, _K_SEMI_MINOR = 2057, 2058
_K_PCS, _K_PROJECTION, _K_TRANSFORM, _K_LINEAR = 3072, 3074, 3075, 3076
_K_STD_PARALLEL1, _K_NAT_LON, _K_NAT_LAT, _K_FALSE_E, _K_FALSE_N = 3078, 3080, 3081, 3082, 3083
_K_FALSE_ORIGIN_E, _K_FALSE_ORIGIN_N = 3090, 3091
_K_CENTER_LON, _K_CENTER_LAT, _K_SCALE_NAT, _K_SCALE_CENTER = 3088, 3089, 3092, 3093
_K_POLE_LON = 3095
_STEREOGRAPHIC_TRANSFORMS = {14, 15, 16}

_TYPE_FORMATS = {1: "B", 2: "s", 3: "H", 4: "I", 5: "II", 6: "b", 8: "h", 9: "i", 10: "ii",
                 11: "f", 12: "d", 16: "Q", 17: "q"}
_TYPE_SIZES = {1: 1, 2: 1, 3: 2, 4: 4, 5: 8, 6: 1, 7: 1, 8: 2, 9: 4, 10: 8, 11: 4, 12: 8,
               16: 8, 17: 8}


class GeoTiffError(LunarscoutError, ValueError):
    """A file that is not a GeoTIFF this reader understands."""


@dataclass(frozen=True)
class GeoTiffInfo:
    """Georeferencing of one band: GDAL-order affine transform, projection, nodata."""

    affine_transform: tuple[float, float, float, float, float, float] | None
    projection: ProjectionParameters | None
    nodata: float | None
    width: int
    height: int


def _dtype(bits: int, sample_format: int) -> np.dtype:
    kind = {1: "u", 2: "i", 3: "f"}.get(sample_format)
    if kind is None or bits not in (8, 16, 32, 64) or (kind == "f" and bits < 32):
        raise GeoTiffError("unsupported TIFF sample type", code="geotiff_data_type_unsupported",
                           details={"bits": bits, "sample_format": sample_format})
    return np.dtype(f"{kind}{bits // 8}")


def _read_ifd(data, endian: str, big: bool = False) -> dict[int, tuple]:
    """Tags of the first IFD of a classic TIFF or a BigTIFF (8-byte offsets)."""
    if big:
        (offset,) = struct.unpack_from(endian + "Q", data, 8)
        (count,) = struct.unpack_from(endian + "Q", data, offset)
        entry_size, head, inline, entry_fmt, ptr_fmt = 20, 8, 8, "HHQ8s", "Q"
    else:
        (offset,) = struct.unpack_from(endian + "I", data, 4)
        (count,) = struct.unpack_from(endian + "H", data, offset)
        entry_size, head, inline, entry_fmt, ptr_fmt = 12, 2, 4, "HHI4s", "I"
    tags: dict[int, tuple] = {}
    for i in range(count):
        tag, typ, n, value = struct.unpack_from(endian + entry_fmt, data, offset + head + entry_size * i)
        size = _TYPE_SIZES.get(typ)
        if size is None:
            continue
        if size * n <= inline:
            raw = bytes(value)
        else:
            start = struct.unpack(endian + ptr_fmt, value)[0]
            raw = bytes(data[start:start + size * n])
        if typ == 2:
            tags[tag] = (raw[:n].split(b"\0")[0].decode("ascii", "replace"),)
        elif typ in (5, 10):
            parts = struct.unpack(endian + _TYPE_FORMATS[typ][0] * (2 * n), raw[: size * n])
            tags[tag] = tuple(parts[2 * k] / parts[2 * k + 1] if parts[2 * k + 1] else 0.0
                              for k in range(n))
        else:
            tags[tag] = struct.unpack(endian + _TYPE_FORMATS[typ] * n, raw[: size * n])
    return tags


def _undo_predictor(block: np.ndarray, predictor: int, dtype: np.dtype, endian: str) -> np.ndarray:
    """block: uint8 [rows, width * itemsize] as stored -> values [rows, width]."""
    rows, row_bytes = block.shape
    size = dtype.itemsize
    width = row_bytes // size
    if predictor == 3:
        planes = np.cumsum(block, axis=1, dtype=np.uint8).reshape(rows, size, width)
        ordered = planes.transpose(0, 2, 1)  # [rows, width, byte] most significant first
        return np.ascontiguousarray(ordered).view(dtype.newbyteorder(">")).reshape(rows, width)
    values = np.ascontiguousarray(block).view(dtype.newbyteorder(endian)).reshape(rows, width)
    if predictor == 2:
        values = np.cumsum(values, axis=1, dtype=values.dtype)
    return values


def _xml(text) -> str:
    return (str(text).replace("&", "&amp;").replace("<", "&lt;").replace(">", "&gt;")
            .replace('"', "&quot;"))


def read_gdal_metadata(path) -> tuple[dict, dict]:
    """GDAL metadata items of a GeoTIFF: ({name: value} of the dataset,
    {band index (0-based): {name: value}}), descriptions included."""
    import xml.etree.ElementTree as ET

    path = Path(path).expanduser()
    with open(path, "rb") as handle:
        head = handle.read(4)
    big = head in (b"II+\0", b"MM\0+")
    endian = "<" if head[:2] == b"II" else ">"
    data = np.memmap(path, dtype=np.uint8, mode="r")
    tags = _read_ifd(data, endian, big)
    del data
    dataset, bands = {}, {}
    if _GDAL_METADATA not in tags:
        return dataset, bands
    for item in ET.fromstring(tags[_GDAL_METADATA][0]).iter("Item"):
        if "sample" in item.attrib:
            bands.setdefault(int(item.attrib["sample"]), {})[item.attrib["name"]] = item.text or ""
        else:
            dataset[item.attrib["name"]] = item.text or ""
    return dataset, bands


def read_geotiff(path, band: int = 1) -> tuple[np.ndarray, GeoTiffInfo]:
    """One band (1-based) of a GeoTIFF or BigTIFF as its native dtype, with its georeferencing."""
    path = Path(path).expanduser()
    if not path.is_file():
        raise GeoTiffError(f"GeoTIFF file does not exist: {path}", code="geotiff_file_not_found",
                           details={"path": str(path)})
    with open(path, "rb") as handle:
        head = handle.read(4)
    big = head in (b"II+\0", b"MM\0+")
    if head in (b"II*\0", b"II+\0"):
        endian = "<"
    elif head in (b"MM\0*", b"MM\0+"):
        endian = ">"
    else:
        raise GeoTiffError(f"File is not a readable GeoTIFF: {path}",
                           code="geotiff_unreadable_or_unsupported", details={"path": str(path)})
    data = np.memmap(path, dtype=np.uint8, mode="r")
    tags = _read_ifd(data, endian, big)
    width, height = int(tags[_WIDTH][0]), int(tags[_HEIGHT][0])
    samples = int(tags.get(_SAMPLES, (1,))[0])
    planar = int(tags.get(_PLANAR, (1,))[0])
    if samples != 1 and planar != 2:
        raise GeoTiffError("only single-band or band-separate TIFFs are supported",
                           code="geotiff_unreadable_or_unsupported", details={"path": str(path)})
    dtype = _dtype(int(tags.get(_BITS, (8,))[0]), int(tags.get(_SAMPLE_FORMAT, (1,))[0]))
    compression = int(tags.get(_COMPRESSION, (1,))[0])
    if compression not in (1, 8, 32946):
        raise GeoTiffError("unsupported TIFF compression", code="geotiff_unreadable_or_unsupported",
                           details={"path": str(path), "compression": compression})
    predictor = int(tags.get(_PREDICTOR, (1,))[0])
    if _TILE_OFFSETS in tags:
        bw, bh = int(tags[_TILE_W][0]), int(tags[_TILE_H][0])
        offsets, counts = tags[_TILE_OFFSETS], tags[_TILE_COUNTS]
    else:
        bw, bh = width, int(tags.get(_ROWS_PER_STRIP, (height,))[0])
        bh = min(bh, height)
        offsets, counts = tags[_STRIP_OFFSETS], tags[_STRIP_COUNTS]
    across = (width + bw - 1) // bw
    down = (height + bh - 1) // bh
    if not 1 <= band <= samples:
        raise GeoTiffError("GeoTIFF band is out of range.", code="geotiff_band_out_of_range",
                           details={"band": band, "band_count": samples})
    first = (band - 1) * across * down  # band-separate blocks are band-major
    out = np.empty((height, width), dtype=dtype)
    for index in range(across * down):
        start, count = offsets[first + index], counts[first + index]
        by, bx = divmod(index, across)
        if count == 0:  # a block never written
            y0, x0 = by * bh, bx * bw
            out[y0:y0 + bh, x0:x0 + bw] = 0
            continue
        raw = bytes(data[start: start + count])
        if compression != 1:
            raw = zlib.decompress(raw)
        rows = bh if _TILE_OFFSETS in tags else min(bh, height - by * bh)
        block = np.frombuffer(raw, dtype=np.uint8)[: rows * bw * dtype.itemsize]
        values = _undo_predictor(block.reshape(rows, bw * dtype.itemsize), predictor, dtype, endian)
        y0, x0 = by * bh, bx * bw
        h, w = min(rows, height - y0), min(bw, width - x0)
        out[y0:y0 + h, x0:x0 + w] = values[:h, :w]
    info = GeoTiffInfo(_affine(tags), _projection(tags), _nodata(tags), width, height)
    del data
    return out, info


def _affine(tags):
    if _TRANSFORMATION in tags:
        m = tags[_TRANSFORMATION]
        return (m[3], m[0], m[1], m[7], m[4], m[5])
    if _TIEPOINT in tags and _PIXEL_SCALE in tags:
        i, j, _k, x, y, _z = tags[_TIEPOINT][:6]
        sx, sy = tags[_PIXEL_SCALE][:2]
        ox, oy = x - i * sx, y + j * sy
        if _geokeys(tags).get(_K_RASTER_TYPE) == 2:  # PixelIsPoint: GDAL shifts by half a pixel
            ox, oy = ox - 0.5 * sx, oy + 0.5 * sy
        return (ox, sx, 0.0, oy, 0.0, -sy)
    return None


def _geokeys(tags) -> dict[int, object]:
    if _GEOKEYS not in tags:
        return {}
    d = tags[_GEOKEYS]
    doubles = tags.get(_GEO_DOUBLES, ())
    ascii_ = tags.get(_GEO_ASCII, ("",))[0]
    keys: dict[int, object] = {}
    for k in range(int(d[3])):
        key, location, count, value = d[4 + 4 * k: 8 + 4 * k]
        if location == 0:
            keys[key] = value
        elif location == _GEO_DOUBLES:
            keys[key] = doubles[value] if count == 1 else tuple(doubles[value:value + count])
        elif location == _GEO_ASCII:
            keys[key] = ascii_[value:value + count].rstrip("|")
    return keys


def _projection(tags) -> ProjectionParameters | None:
    keys = _geokeys(tags)
    if not keys:
        return None
    transform = keys.get(_K_TRANSFORM)
    if transform not in _STEREOGRAPHIC_TRANSFORMS:
        raise GeoTiffError("Horizon-derived products require a stereographic lunar DEM.",
                           code="product_dem_projection_unsupported",
                           details={"coordinate_transformation": transform})

    def first(*names):
        for n in names:
            if n in keys:
                return float(keys[n])
        return None

    lat0 = first(_K_NAT_LAT, _K_CENTER_LAT)
    std = first(_K_STD_PARALLEL1)
    if lat0 is None and std is not None and transform == 15:
        lat0 = math.copysign(90.0, std)
    lon0 = first(_K_POLE_LON, _K_NAT_LON, _K_CENTER_LON)
    scale = first(_K_SCALE_NAT, _K_SCALE_CENTER)
    if scale is None and std is not None and abs(abs(std) - 90.0) <= 1e-10:
        scale = 1.0
    fe = first(_K_FALSE_E, _K_FALSE_ORIGIN_E)
    fn = first(_K_FALSE_N, _K_FALSE_ORIGIN_N)
    radius = first(_K_SEMI_MAJOR)
    values = (radius, lat0, lon0, scale, fe, fn)
    if any(v is None or not math.isfinite(v) for v in values):
        raise GeoTiffError("The DEM stereographic projection parameters are incomplete.",
                           code="product_dem_projection_invalid", details={})
    return ProjectionParameters(radius, math.radians(lat0), math.radians(lon0), scale, fe, fn)


def _nodata(tags) -> float | None:
    if _GDAL_NODATA not in tags:
        return None
    try:
        return float(tags[_GDAL_NODATA][0].strip())
    except ValueError:
        return None


def load_dem(path) -> DemGrid:
    """A stereographic DEM as ``DemGrid`` (``products.py:97-177``)."""
    dem_path = Path(path).expanduser().resolve()
    if not dem_path.is_file():
        raise InputError("The DEM path does not identify a file.", code="product_dem_not_found",
                         details={"path": str(dem_path)})
    values, info = read_geotiff(dem_path)
    if info.affine_transform is None or info.projection is None:
        raise GeoTiffError("The DEM must have complete geospatial metadata.",
                           code="product_dem_not_georeferenced", details={"path": str(dem_path)})
    return DemGrid(np.ascontiguousarray(values, dtype=np.float32),
                   np.asarray(info.affine_transform, dtype=np.float64), info.projection)


# -- writer -----------------------------------------------------------------------------------

_TIFF_TYPES = {3: ("H", 2), 4: ("I", 4), 12: ("d", 8), 16: ("Q", 8), 2: ("s", 1)}


def _encode_block(values: np.ndarray, predictor: int, compress: bool) -> bytes:
    rows, width = values.shape
    if not compress:
        return np.ascontiguousarray(values.astype(values.dtype.newbyteorder("<"))).tobytes()
    if predictor == 3:
        be = np.ascontiguousarray(values.astype(values.dtype.newbyteorder(">")))
        planes = be.view(np.uint8).reshape(rows, width, values.dtype.itemsize).transpose(0, 2, 1)
        planes = np.ascontiguousarray(planes).reshape(rows, -1)
        diff = np.diff(planes, axis=1, prepend=np.zeros((rows, 1), np.uint8)).astype(np.uint8)
        return zlib.compress(diff.tobytes(), 6)
    le = np.ascontiguousarray(values.astype(values.dtype.newbyteorder("<")))
    diff = np.diff(le, axis=1, prepend=np.zeros((rows, 1), le.dtype)).astype(le.dtype)
    return zlib.compress(diff.tobytes(), 6)


class GeoTiffWriter:
    """Streaming writer of a tiled (128x128), band-separate GeoTIFF (geotiff.py:190-260).

    Blocks may be written in any order (``write_block`` / ``write_region``);
    ``close`` appends the directory and atomically renames the staged file over
    ``path``. Deflate with predictor 3 (floats) or 2 (integers), as the
    reference writes its products; a BigTIFF (8-byte offsets) when the raw size
    could exceed 4 GiB. Blocks never written stay empty (read as zeros).
    """

    TILE = 128

    def __init__(self, path, width: int, height: int, bands: int, dtype, affine_transform,
                 projection: ProjectionParameters, *, nodata: float | None = None,
                 overwrite: bool = False, compress: bool = True,
                 descriptions: list[str] | None = None, staging=None, resume=None,
                 metadata: dict | None = None, band_metadata: list | None = None):
        """``staging``: a fixed staging file instead of a temporary one (the
        resumable product store); ``resume``: (offsets, counts, end) of the
        blocks already in that file, which is truncated to ``end`` and appended to.
        ``metadata`` / ``band_metadata`` (one dict per band): GDAL metadata items
        of the dataset and of each band, written in the GDAL_METADATA tag as GDAL
        keeps ``update_tags`` items (the reference tags its products so)."""
        self.path = Path(path).expanduser().resolve()
        self.dtype = np.dtype(dtype)
        if self.dtype == np.bool_ or self.dtype.kind not in "uif":
            raise InputError("unsupported GeoTIFF data type", code="geotiff_data_type_unsupported",
                             details={"dtype": str(self.dtype)})
        if self.path.exists() and not overwrite:
            raise InputError(f"GeoTIFF output already exists: {self.path}", code="output_exists",
                             details={"path": str(self.path)})
        self.width, self.height, self.bands = int(width), int(height), int(bands)
        self.affine = tuple(float(v) for v in affine_transform)
        self.projection = projection
        self.nodata = nodata
        self.compress = bool(compress)
        self.predictor = (3 if self.dtype.kind == "f" else 2) if self.compress else 1
        self.across = (self.width + self.TILE - 1) // self.TILE
        self.down = (self.height + self.TILE - 1) // self.TILE
        blocks = self.across * self.down * self.bands
        raw = blocks * self.TILE * self.TILE * self.dtype.itemsize
        self.big = raw + 16 * blocks + (1 << 20) >= (1 << 32)
        self.offsets = np.zeros(blocks, dtype=np.uint64)
        self.counts = np.zeros(blocks, dtype=np.uint64)
        self.descriptions = descriptions
        self.metadata = {str(k): str(v) for k, v in (metadata or {}).items()}
        self.band_metadata = [{str(k): str(v) for k, v in (m or {}).items()}
                              for m in (band_metadata or [])]
        self.path.parent.mkdir(parents=True, exist_ok=True)
        if staging is None:
            fd, tmp = tempfile.mkstemp(prefix=f".{self.path.name}.", suffix=".tmp.tif",
                                       dir=self.path.parent)
            self.tmp = Path(tmp)
            self.handle = os.fdopen(fd, "wb")
        else:
            self.tmp = Path(staging)
            self.handle = open(self.tmp, "r+b" if resume is not None else "w+b")
        if resume is None:
            self.handle.write(b"\0" * (16 if self.big else 8))  # header, patched on close
        else:
            offsets, counts, end = resume
            self.offsets[:] = np.asarray(offsets, dtype=np.uint64)
            self.counts[:] = np.asarray(counts, dtype=np.uint64)
            self.handle.truncate(int(end))
            self.handle.seek(int(end))

    def encode_block(self, block: np.ndarray) -> bytes:
        """The stored bytes of one block (smaller edge blocks are zero-padded to
        128x128). Touches no writer state, so blocks may be encoded on several
        threads (zlib releases the GIL) and written in order afterwards."""
        tile = np.zeros((self.TILE, self.TILE), dtype=self.dtype)
        part = np.asarray(block, dtype=self.dtype)
        tile[: part.shape[0], : part.shape[1]] = part
        return _encode_block(tile, self.predictor, self.compress)

    def write_encoded(self, band: int, by: int, bx: int, payload: bytes) -> None:
        """Append a block's stored bytes (``encode_block``) and record its place."""
        index = (band * self.down + by) * self.across + bx
        self.offsets[index] = self.handle.tell()
        self.counts[index] = len(payload)
        self.handle.write(payload)
        if len(payload) % 2:
            self.handle.write(b"\0")

    def write_block(self, band: int, by: int, bx: int, block: np.ndarray) -> None:
        """One 128x128 block (smaller edge blocks are zero-padded) of a 0-based band."""
        self.write_encoded(band, by, bx, self.encode_block(block))

    def block_index(self, band: int, by: int, bx: int) -> int:
        return (band * self.down + by) * self.across + bx

    def sync(self) -> int:
        """Flush the staged data to disk; returns the end offset."""
        self.handle.flush()
        os.fsync(self.handle.fileno())
        return self.handle.tell()

    def write_region(self, band: int, y0: int, x0: int, values: np.ndarray) -> None:
        """A block-aligned region (y0, x0 multiples of 128) of a 0-based band."""
        if y0 % self.TILE or x0 % self.TILE:
            raise ValueError("regions must start on the 128-pixel block grid")
        h, w = values.shape
        for by in range(y0 // self.TILE, (y0 + h + self.TILE - 1) // self.TILE):
            for bx in range(x0 // self.TILE, (x0 + w + self.TILE - 1) // self.TILE):
                ry, rx = by * self.TILE - y0, bx * self.TILE - x0
                self.write_block(band, by, bx, values[ry: ry + self.TILE, rx: rx + self.TILE])

    def _tags(self):
        t = self.affine
        geokeys, doubles = _stereographic_geokeys(self.projection)
        fmt = {"u": 1, "i": 2, "f": 3}[self.dtype.kind]
        offset_type = 16 if self.big else 4
        tags = [
            (_WIDTH, 4, (self.width,)), (_HEIGHT, 4, (self.height,)),
            (_BITS, 3, (self.dtype.itemsize * 8,) * self.bands),
            (_COMPRESSION, 3, (8 if self.compress else 1,)), (_PHOTOMETRIC, 3, (1,)),
            (_SAMPLES, 3, (self.bands,)), (_PLANAR, 3, (2 if self.bands > 1 else 1,)),
            (_PREDICTOR, 3, (self.predictor,)), (_TILE_W, 3, (self.TILE,)),
            (_TILE_H, 3, (self.TILE,)),
            (_TILE_OFFSETS, offset_type, tuple(int(v) for v in self.offsets)),
            (_TILE_COUNTS, offset_type, tuple(int(v) for v in self.counts)),
            (_SAMPLE_FORMAT, 3, (fmt,) * self.bands),
            (_GEOKEYS, 3, geokeys), (_GEO_DOUBLES, 12, doubles),
        ]
        if t[2] != 0.0 or t[4] != 0.0:
            tags.append((_TRANSFORMATION, 12, (t[1], t[2], 0.0, t[0], t[4], t[5], 0.0, t[3],
                                               0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)))
        else:
            tags.append((_PIXEL_SCALE, 12, (t[1], -t[5], 0.0)))
            tags.append((_TIEPOINT, 12, (0.0, 0.0, 0.0, t[0], t[3], 0.0)))
        if self.nodata is not None:
            tags.append((_GDAL_NODATA, 2, repr(float(self.nodata))))
        items = [f'<Item name="{_xml(k)}">{_xml(v)}</Item>' for k, v in self.metadata.items()]
        for i, m in enumerate(self.band_metadata):
            items += [f'<Item name="{_xml(k)}" sample="{i}">{_xml(v)}</Item>' for k, v in m.items()]
        for i, d in enumerate(self.descriptions or []):
            items.append(f'<Item name="DESCRIPTION" sample="{i}" role="description">{_xml(d)}</Item>')
        if items:
            tags.append((_GDAL_METADATA, 2, "<GDALMetadata>" + "".join(items) + "</GDALMetadata>"))
        return sorted(tags, key=lambda x: x[0])

    def close(self) -> Path:
        e = "<"
        big = self.big
        tags = self._tags()
        entry_size, count_fmt, inline = (20, "Q", 8) if big else (12, "H", 4)
        # out-of-line values first, then the directory
        values_at = {}
        for tag, typ, vals in tags:
            code, size = _TIFF_TYPES[typ]
            raw = (vals.encode("ascii") + b"\0") if typ == 2 else struct.pack(e + code * len(vals), *vals)
            if len(raw) > inline:
                values_at[tag] = self.handle.tell()
                self.handle.write(raw)
                if len(raw) % 2:
                    self.handle.write(b"\0")
        ifd = self.handle.tell()
        self.handle.write(struct.pack(e + count_fmt, len(tags)))
        for tag, typ, vals in tags:
            code, size = _TIFF_TYPES[typ]
            raw = (vals.encode("ascii") + b"\0") if typ == 2 else struct.pack(e + code * len(vals), *vals)
            n = len(raw) if typ == 2 else len(vals)
            if big:
                head = struct.pack(e + "HHQ", tag, typ, n)
                body = raw.ljust(8, b"\0") if len(raw) <= 8 else struct.pack(e + "Q", values_at[tag])
            else:
                head = struct.pack(e + "HHI", tag, typ, n)
                body = raw.ljust(4, b"\0") if len(raw) <= 4 else struct.pack(e + "I", values_at[tag])
            self.handle.write(head + body)
        self.handle.write(struct.pack(e + ("Q" if big else "I"), 0))
        self.handle.seek(0)
        if big:
            self.handle.write(b"II+\0" + struct.pack(e + "HHQ", 8, 0, ifd))
        else:
            self.handle.write(b"II*\0" + struct.pack(e + "I", ifd))
        self.handle.flush()
        os.fsync(self.handle.fileno())
        self.handle.close()
        os.replace(self.tmp, self.path)
        return self.path

    def abort(self) -> None:
        try:
            self.handle.close()
        finally:
            self.tmp.unlink(missing_ok=True)

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc, tb):
        if exc_type is None:
            self.close()
        else:
            self.abort()


def write_geotiff(path, array: np.ndarray, affine_transform, projection: ProjectionParameters, *,
                  nodata: float | None = None, overwrite: bool = False) -> Path:
    """Atomically write one band as a tiled 128x128 deflate GeoTIFF (geotiff.py:202-260)."""
    values = np.asarray(array)
    if values.ndim != 2:
        raise InputError("GeoTIFF output must be a two-dimensional array.",
                         code="geotiff_invalid_array_shape", details={"shape": list(values.shape)})
    writer = GeoTiffWriter(path, values.shape[1], values.shape[0], 1, values.dtype,
                           affine_transform, projection, nodata=nodata, overwrite=overwrite)
    try:
        writer.write_region(0, 0, 0, values)
    except BaseException:
        writer.abort()
        raise
    return writer.close()


def _stereographic_geokeys(p: ProjectionParameters):
    lat0 = math.degrees(p.latitude_origin_rad)
    polar = abs(abs(lat0) - 90.0) <= 1e-10
    doubles = [lat0, math.degrees(p.longitude_origin_rad), p.scale, p.false_easting_m,
               p.false_northing_m, p.radius_m, p.radius_m]
    lon_key = _K_POLE_LON if polar else _K_NAT_LON
    keys = [
        (_K_MODEL_TYPE, 0, 1, 1), (_K_RASTER_TYPE, 0, 1, 1),
        (_K_GEOG_TYPE, 0, 1, 32767), (_K_DATUM, 0, 1, 32767), (_K_ANGULAR, 0, 1, 9102),
        (_K_ELLIPSOID, 0, 1, 32767), (_K_SEMI_MAJOR, _GEO_DOUBLES, 1, 5),
        (_K_SEMI_MINOR, _GEO_DOUBLES, 1, 6),
        (_K_PCS, 0, 1, 32767), (_K_PROJECTION, 0, 1, 32767),
        (_K_TRANSFORM, 0, 1, 15 if polar else 14), (_K_LINEAR, 0, 1, 9001),
        (_K_NAT_LAT, _GEO_DOUBLES, 1, 0), (lon_key, _GEO_DOUBLES, 1, 1),
        (_K_SCALE_NAT, _GEO_DOUBLES, 1, 2), (_K_FALSE_E, _GEO_DOUBLES, 1, 3),
        (_K_FALSE_N, _GEO_DOUBLES, 1, 4),
    ]
    keys.sort(key=lambda k: k[0])
    flat = [1, 1, 0, len(keys)]
    for k in keys:
        flat.extend(k)
    return tuple(flat), tuple(doubles)


def write_dem(path, dem: DemGrid, *, overwrite: bool = False) -> Path:
    """A ``DemGrid`` as a float32 GeoTIFF that ``load_dem`` reads back exactly."""
    return write_geotiff(path, np.asarray(dem.elevation_m, dtype=np.float32),
                         tuple(float(v) for v in dem.geo_transform), dem.projection,
                         overwrite=overwrite)
