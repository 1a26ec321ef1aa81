"""Resumable tiled GeoTIFF products: staged file, job manifest, patch journal.

Counterpart of the reference's ``ResumableTiledProduct``
(src/lunarscout/_numba_horizon/product_store.py:154-653) without rasterio:

* the product is written to ``.<name>.lunarscout-partial.tif`` beside the
  output, with ``.manifest.json`` (the job identity: size, dtype, bands,
  nodata, compression, georeferencing, algorithm and configuration, and its
  SHA-256 fingerprint) and ``.journal.jsonl`` beside it;
* every finished patch appends one journal line -- its key, valid / invalid,
  and the (offset, byte count) of each band block it wrote -- after the blocks
  are on disk (fsync), so the journal never names data that is not there;
* a run that finds all three files resumes: a different job identity raises
  ``IncompatibleProductJobError`` (``start_fresh=True`` discards the staged
  product); the staged file is cut back to the end of the last journaled
  patch (a torn last line, or blocks written after it, are dropped) and the
  completed patches are skipped;
* ``finalize`` appends the TIFF directory, renames the staged file over the
  output and removes the manifest and journal.
"""

from __future__ import annotations

import hashlib
import json
import os
from pathlib import Path
from typing import Any

import numpy as np

from .errors import ProductGenerationError
from .geometry import projection_array
from .geotiff import GeoTiffWriter

MANIFEST_SCHEMA = "lunarscout-amd-product-job-v1"
PATCH_SIZE = 128


class IncompatibleProductJobError(ProductGenerationError):
    """A staged product exists for a different job (product_store.py:54-55)."""


def _json_bytes(value: Any) -> bytes:
    return json.dumps(value, sort_keys=True, separators=(",", ":"), allow_nan=False).encode()


def _atomic_json(path: Path, value: Any) -> None:
    tmp = path.with_name(path.name + ".tmp")
    with open(tmp, "wb") as handle:
        handle.write(_json_bytes(value))
        handle.flush()
        os.fsync(handle.fileno())
    os.replace(tmp, path)


def _nodata_json(value):
    if value is None:
        return None
    value = float(value)
    return "NaN" if np.isnan(value) else value


class ResumableTiledProduct:
    """A tiled, band-separate product written patch by patch, resumable."""

    def __init__(self, output_path, *, width: int, height: int, bands: int, dtype, affine_transform,
                 projection, nodata=None, compress: bool = True, descriptions=None,
                 algorithm: str = "unspecified", configuration: dict | None = None,
                 overwrite: bool = False, start_fresh: bool = False):
        self.output_path = Path(output_path).expanduser().resolve()
        self.staging_path = self.output_path.with_name(
            f".{self.output_path.name}.lunarscout-partial.tif")
        self.manifest_path = self.staging_path.with_suffix(".manifest.json")
        self.journal_path = self.staging_path.with_suffix(".journal.jsonl")
        if self.output_path.exists() and not overwrite:
            raise ProductGenerationError("The product output already exists.",
                                         code="product_output_exists",
                                         details={"path": str(self.output_path)})
        manifest = {
            "schema": MANIFEST_SCHEMA, "algorithm": str(algorithm),
            "width": int(width), "height": int(height), "bands": int(bands),
            "dtype": np.dtype(dtype).str, "nodata": _nodata_json(nodata),
            "compression": "deflate" if compress else "none",
            "affine_transform": [float(v) for v in affine_transform],
            "projection": [float(v) for v in projection_array(projection)],
            "descriptions": list(descriptions or []),
            "configuration": configuration or {},
            "tile_size": PATCH_SIZE, "interleave": "band",
        }
        self.fingerprint = hashlib.sha256(_json_bytes(manifest)).hexdigest()
        manifest["fingerprint"] = self.fingerprint
        self.completed: dict[str, str] = {}
        if start_fresh:
            self._remove_staging()
        present = [p.exists() for p in (self.staging_path, self.manifest_path, self.journal_path)]
        if any(present) and not all(present):
            raise IncompatibleProductJobError(
                "The staged product is incomplete; use start_fresh=True to discard it.",
                code="product_staging_incomplete", details={"path": str(self.staging_path)})
        resume = None
        if all(present):
            stored = json.loads(self.manifest_path.read_text())
            if stored.get("fingerprint") != self.fingerprint:
                raise IncompatibleProductJobError(
                    "A staged product of a different job exists; use start_fresh=True.",
                    code="product_job_incompatible", details={"path": str(self.staging_path)})
            resume = self._read_journal(width, height, bands)
        self.writer = GeoTiffWriter(self.output_path, width, height, bands, dtype,
                                    affine_transform, projection, nodata=nodata,
                                    overwrite=True, compress=compress, descriptions=descriptions,
                                    staging=self.staging_path, resume=resume)
        if resume is None:
            self.completed = {}
            self.writer.sync()
            _atomic_json(self.manifest_path, manifest)
            with open(self.journal_path, "wb") as handle:
                handle.flush()
                os.fsync(handle.fileno())
        self._journal = open(self.journal_path, "ab")

    # -- resume ---------------------------------------------------------------------
    def _read_journal(self, width, height, bands):
        """(offsets, counts, end) of the journaled patches, or None when there
        are none (the staged product then starts over)."""
        across = (width + PATCH_SIZE - 1) // PATCH_SIZE
        down = (height + PATCH_SIZE - 1) // PATCH_SIZE
        offsets = np.zeros(across * down * bands, dtype=np.uint64)
        counts = np.zeros_like(offsets)
        end = 0
        good = 0
        raw = self.journal_path.read_bytes()
        for line in raw.splitlines(keepends=True):
            if not line.endswith(b"\n"):
                break  # a torn last line: that patch is written again
            entry = json.loads(line)
            for index, offset, count in entry["blocks"]:
                offsets[index], counts[index] = offset, count
            self.completed[entry["patch"]] = entry["state"]
            end = max(end, int(entry["end"]))
            good += len(line)
        if good != len(raw):
            with open(self.journal_path, "r+b") as handle:
                handle.truncate(good)
        if not self.completed:
            return None
        return offsets, counts, end

    def _remove_staging(self) -> None:
        for path in (self.staging_path, self.manifest_path, self.journal_path):
            path.unlink(missing_ok=True)

    # -- writing ----------------------------------------------------------------------
    @staticmethod
    def key(tile_y: int, tile_x: int) -> str:
        return f"{tile_y},{tile_x}"

    def is_complete(self, tile_y: int, tile_x: int) -> bool:
        return self.key(tile_y, tile_x) in self.completed

    def write_patch(self, tile_y: int, tile_x: int, blocks, *, state: str = "valid") -> None:
        """All bands of one patch (``blocks``: iterable of [h, w] arrays in band
        order), then its journal line."""
        by, bx = tile_y // PATCH_SIZE, tile_x // PATCH_SIZE
        written = []
        for band, block in enumerate(blocks):
            index = self.writer.block_index(band, by, bx)
            self.writer.write_block(band, by, bx, block)
            written.append([index, int(self.writer.offsets[index]), int(self.writer.counts[index])])
        if len(written) != self.writer.bands:
            raise ValueError(f"a patch needs {self.writer.bands} bands, got {len(written)}")
        end = self.writer.sync()
        key = self.key(tile_y, tile_x)
        self._journal.write(_json_bytes({"patch": key, "state": state, "end": end,
                                         "blocks": written}) + b"\n")
        self._journal.flush()
        os.fsync(self._journal.fileno())
        self.completed[key] = state

    def write_invalid_patch(self, tile_y: int, tile_x: int, height: int, width: int,
                            fill) -> None:
        block = np.full((height, width), fill, dtype=self.writer.dtype)
        self.write_patch(tile_y, tile_x, (block for _ in range(self.writer.bands)),
                         state="invalid")

    def finalize(self) -> Path:
        self._journal.close()
        path = self.writer.close()
        self.manifest_path.unlink(missing_ok=True)
        self.journal_path.unlink(missing_ok=True)
        return path

    def suspend(self) -> None:
        """Stop writing and keep the staged product for a later resume."""
        self._journal.close()
        self.writer.handle.close()
