"""Resumable tiled GeoTIFF products: staged file, job manifest, patch journal.

Counterpart of the reference's ``ResumableTiledProduct`` and ``ProductJob``
(src/lunarscout/_numba_horizon/product_store.py:154-653) without rasterio:

* the product is written to ``.<name>.lunarscout-partial.tif`` beside the
  output, with ``.manifest.json`` and ``.journal.jsonl`` beside it;
* the manifest holds the job identity with the reference's fields
  (product_store.py:154-252): algorithm and version, size, dtype, band count,
  the UTC timestamp of every band, per-band metadata, the invalid value,
  nodata, compression, georeferencing, the configuration and the identity of
  the horizon tiles the job reads (``horizon_inventory_identity``: path, size
  and modification time of every patch's tile), and its SHA-256 fingerprint;
  the backends that computed valid patches follow it as ``compute_backends``;
* every finished patch appends one journal line -- its key, valid / invalid,
  its backend and the (offset, byte count) of each band block it wrote --
  after the blocks are on disk (fsync), so the journal never names data that
  is not there. ``write_patch`` makes each patch durable by itself;
  ``batch_writer(n)`` writes patches back to back and makes them durable n at a
  time (one fsync of the staged file and one of the journal per n patches --
  the reference's ``ProductBatchWriter``, product_store.py:655-794), encoding
  the band blocks on a thread pool; a crash loses at most the patches since
  the last checkpoint, which a resume computes again;
* ``output_transform`` (product_store.py:362-385) maps every valid band block
  before it is stored and must keep its shape and return the product's dtype;
* a run that finds all three files resumes: a different job identity (a
  different horizon store included) raises ``IncompatibleProductJobError``
  (``start_fresh=True`` discards the staged product); the staged file is cut
  back to the end of the last journaled patch (a torn last line, or blocks
  written after it, are dropped) and the completed patches are skipped;
* ``finalize`` appends the TIFF directory with the reference's GDAL metadata
  tags (LUNARSCOUT_PRODUCT_SCHEMA, LUNARSCOUT_TIMESTAMPS_UTC,
  LUNARSCOUT_COMPUTE_BACKENDS on the dataset, TIMESTAMP_UTC and the band
  metadata on each band; product_store.py:27-32, 485-510, 624-640), renames the
  staged file over the output and removes the manifest and journal.
"""

from __future__ import annotations

import hashlib
import json
import os
from concurrent.futures import ThreadPoolExecutor
from datetime import datetime
from pathlib import Path
from typing import Any, Callable

import numpy as np

from .errors import ProductGenerationError
from .geometry import projection_array
from .geotiff import GeoTiffWriter

MANIFEST_SCHEMA = "lunarscout-amd-product-job-v2"
ALGORITHM_VERSION = "lunarscout-amd-lighting-v2"
PATCH_SIZE = 128
SCHEMA_TAG = "LUNARSCOUT_PRODUCT_SCHEMA"
TIMESTAMPS_TAG = "LUNARSCOUT_TIMESTAMPS_UTC"
TIMESTAMP_TAG = "TIMESTAMP_UTC"
COMPUTE_BACKENDS_TAG = "LUNARSCOUT_COMPUTE_BACKENDS"


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


def utc_stamp(value) -> str:
    """A band timestamp as the manifest and the tags hold it: UTC, microseconds, 'Z'."""
    if isinstance(value, datetime):
        stamp = value
    else:
        text = str(value).strip()
        stamp = datetime.fromisoformat(text[:-1] + "+00:00" if text.endswith("Z") else text)
    offset = stamp.utcoffset()
    if offset is None:
        raise ValueError("timestamps must include a UTC offset")
    utc = (stamp - offset).replace(tzinfo=None)
    return utc.strftime("%Y-%m-%dT%H:%M:%S.%fZ")


def horizon_inventory_identity(store, patches, observer_elevation_m: float) -> str:
    """Identity of the horizon tiles a product job reads.

    One record per patch -- its tile's resolved path, byte size and
    modification time, or a marker when the tile is missing -- hashed, so a
    tile regenerated (or added, or removed) between a cancelled run and its
    resume changes the job's fingerprint without re-reading any payload
    (psr_pipeline.py:96-123 keys the reference's jobs the same way).
    """
    digest = hashlib.sha256()
    for patch in patches:
        path = store.find_existing_path(patch.tile_y, patch.tile_x, observer_elevation_m)
        if path is None:
            record = f"{patch.tile_y},{patch.tile_x},absent"
        else:
            st = Path(path).stat()
            record = f"{patch.tile_y},{patch.tile_x},{Path(path).resolve()},{st.st_size},{st.st_mtime_ns}"
        digest.update(record.encode("utf-8") + b"\n")
    return "stat-sha256:" + digest.hexdigest()


class ResumableTiledProduct:
    """A tiled, band-separate product written patch by patch, resumable."""

    def __init__(self, output_path, *, width: int, height: int, bands: int, dtype, affine_transform,
                 projection, nodata=None, compress: bool = True, descriptions=None,
                 algorithm: str = "unspecified", configuration: dict | None = None,
                 timestamps_utc=(), band_metadata=(), invalid_value=0,
                 horizon_inventory_identity: str = "unspecified", backend: str | None = None,
                 overwrite: bool = False, start_fresh: bool = False,
                 output_transform: Callable[[np.ndarray], np.ndarray] | None = None):
        self.output_path = Path(output_path).expanduser().resolve()
        self.staging_path = self.output_path.with_name(
            f".{self.output_path.name}.lunarscout-partial.tif")
        self.manifest_path = self.staging_path.with_suffix(".manifest.json")
        self.journal_path = self.staging_path.with_suffix(".journal.jsonl")
        if self.output_path.exists() and not overwrite:
            raise ProductGenerationError("The product output already exists.",
                                         code="product_output_exists",
                                         details={"path": str(self.output_path)})
        timestamps = [utc_stamp(t) for t in timestamps_utc]
        if timestamps and len(timestamps) != int(bands):
            raise ValueError("timestamp count must equal the band count")
        band_metadata = [dict(m) for m in band_metadata] or [{} for _ in range(int(bands))]
        if len(band_metadata) != int(bands):
            raise ValueError("band metadata count must equal the band count")
        manifest = {
            "schema": MANIFEST_SCHEMA, "algorithm_version": ALGORITHM_VERSION,
            "algorithm": str(algorithm),
            "width": int(width), "height": int(height), "dtype": np.dtype(dtype).str,
            "band_count": int(bands), "timestamps_utc": timestamps,
            "band_metadata": band_metadata, "invalid_value": _nodata_json(invalid_value),
            "nodata": _nodata_json(nodata),
            "compression": "deflate" if compress else "none",
            "affine_transform": [float(v) for v in affine_transform],
            "projection": [float(v) for v in projection_array(projection)],
            "descriptions": list(descriptions or []),
            "configuration": configuration or {},
            "horizon_inventory_identity": str(horizon_inventory_identity),
            "tile_size": PATCH_SIZE, "interleave": "band",
        }
        self.fingerprint = hashlib.sha256(_json_bytes(manifest)).hexdigest()
        manifest["fingerprint"] = self.fingerprint
        self._manifest = manifest
        self.output_transform = output_transform
        self.backend = backend
        self.compute_backends: list[str] = []
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
        per_band = [{str(k): str(v) for k, v in m.items()} for m in band_metadata]
        for i, stamp in enumerate(timestamps):
            per_band[i][TIMESTAMP_TAG] = stamp
        self.writer = GeoTiffWriter(self.output_path, width, height, bands, dtype,
                                    affine_transform, projection, nodata=nodata,
                                    overwrite=True, compress=compress, descriptions=descriptions,
                                    staging=self.staging_path, resume=resume,
                                    metadata={SCHEMA_TAG: MANIFEST_SCHEMA,
                                              TIMESTAMPS_TAG: json.dumps(timestamps,
                                                                         separators=(",", ":"))},
                                    band_metadata=per_band)
        if resume is None:
            self.completed = {}
            self.compute_backends = []
            self.writer.sync()
            self._write_manifest()
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
            backend = entry.get("backend")
            if entry["state"] == "valid" and backend and backend not in self.compute_backends:
                self.compute_backends.append(backend)
            end = max(end, int(entry["end"]))
            good += len(line)
        if good != len(raw):
            with open(self.journal_path, "r+b") as handle:
                handle.truncate(good)
        if not self.completed:
            return None
        return offsets, counts, end

    def _write_manifest(self) -> None:
        _atomic_json(self.manifest_path, dict(self._manifest,
                                              compute_backends=list(self.compute_backends)))

    def _remove_staging(self) -> None:
        for path in (self.staging_path, self.manifest_path, self.journal_path):
            path.unlink(missing_ok=True)

    # -- writing ----------------------------------------------------------------------
    @staticmethod
    def key(tile_y: int, tile_x: int) -> str:
        return f"{tile_y},{tile_x}"

    def is_complete(self, tile_y: int, tile_x: int) -> bool:
        return self.key(tile_y, tile_x) in self.completed

    def _prepare(self, block, state: str) -> np.ndarray:
        block = np.asarray(block)
        if state != "valid" or self.output_transform is None:
            return block
        mapped = np.asarray(self.output_transform(block))
        if mapped.shape != block.shape:
            raise ValueError(f"output_transform must keep the block shape {block.shape}, "
                             f"got {mapped.shape}")
        if mapped.dtype != self.writer.dtype:
            raise ValueError(f"output_transform returned {mapped.dtype}; the product holds "
                             f"{self.writer.dtype}")
        return mapped

    def _store_patch(self, tile_y: int, tile_x: int, blocks, state: str, encode=None) -> dict:
        """Write every band block of a patch; return its (not yet durable) journal entry."""
        key = self.key(tile_y, tile_x)
        if key in self.completed:
            raise ValueError(f"patch {key} is already complete")
        by, bx = tile_y // PATCH_SIZE, tile_x // PATCH_SIZE
        prepared = [self._prepare(block, state) for block in blocks]
        if len(prepared) != self.writer.bands:
            raise ValueError(f"a patch needs {self.writer.bands} bands, got {len(prepared)}")
        if encode is None:
            payloads = [self.writer.encode_block(b) for b in prepared]
        else:
            payloads = list(encode(self.writer.encode_block, prepared))
        written = []
        for band, payload in enumerate(payloads):
            index = self.writer.block_index(band, by, bx)
            self.writer.write_encoded(band, by, bx, payload)
            written.append([index, int(self.writer.offsets[index]), int(self.writer.counts[index])])
        entry = {"patch": key, "state": state, "blocks": written}
        if state == "valid" and self.backend:
            entry["backend"] = self.backend
        return entry

    def _checkpoint(self, entries: list[dict]) -> list[str]:
        """Make the written patches durable: the staged blocks first, then their
        journal lines. Returns their keys."""
        if not entries:
            return []
        end = self.writer.sync()
        fresh = [e["backend"] for e in entries
                 if e.get("backend") and e["backend"] not in self.compute_backends]
        for backend in fresh:
            if backend not in self.compute_backends:
                self.compute_backends.append(backend)
        if fresh:
            self._write_manifest()
        self._journal.write(b"".join(_json_bytes(dict(e, end=end)) + b"\n" for e in entries))
        self._journal.flush()
        os.fsync(self._journal.fileno())
        for e in entries:
            self.completed[e["patch"]] = e["state"]
        return [e["patch"] for e in entries]

    def write_patch(self, tile_y: int, tile_x: int, blocks, *, state: str = "valid") -> None:
        """All bands of one patch (``blocks``: iterable of [h, w] arrays in band
        order), made durable with its journal line before returning."""
        self._checkpoint([self._store_patch(tile_y, tile_x, blocks, state)])

    def write_invalid_patch(self, tile_y: int, tile_x: int, height: int, width: int,
                            fill) -> None:
        block = np.full((height, width), fill, dtype=self.writer.dtype)
        self.write_patch(tile_y, tile_x, [block] * self.writer.bands, state="invalid")

    def batch_writer(self, checkpoint_patch_count: int, *,
                     encode_workers: int = 0) -> "ProductBatchWriter":
        """A writer that checkpoints every ``checkpoint_patch_count`` patches."""
        return ProductBatchWriter(self, checkpoint_patch_count, encode_workers=encode_workers)

    def finalize(self) -> Path:
        self._journal.close()
        self.writer.metadata[COMPUTE_BACKENDS_TAG] = json.dumps(self.compute_backends,
                                                                separators=(",", ":"))
        path = self.writer.close()
        self.manifest_path.unlink(missing_ok=True)
        self.journal_path.unlink(missing_ok=True)
        return path

    def suspend(self) -> None:
        """Stop writing and keep the staged product for a later resume."""
        self._journal.close()
        self.writer.handle.close()


class ProductBatchWriter:
    """Patches written back to back, made durable ``checkpoint_patch_count`` at a
    time (product_store.py:655-794). ``write_patch`` and ``checkpoint`` return
    the keys of the patches a checkpoint made durable (empty when none did);
    leaving the ``with`` block checkpoints what was written, also on an error
    (every written patch is whole: a patch enters the batch after its last
    band). ``encode_workers`` > 1 compresses a patch's band blocks on that many
    threads."""

    def __init__(self, product: ResumableTiledProduct, checkpoint_patch_count: int, *,
                 encode_workers: int = 0):
        if isinstance(checkpoint_patch_count, bool) or not isinstance(checkpoint_patch_count, int) \
                or checkpoint_patch_count < 1:
            raise ValueError("checkpoint_patch_count must be a positive integer")
        self.product = product
        self.limit = checkpoint_patch_count
        self.pending: list[dict] = []
        self.pool = ThreadPoolExecutor(encode_workers, thread_name_prefix="product-encode") \
            if encode_workers > 1 else None
        self.invalid_payloads: dict[tuple, bytes] = {}
        self.closed = False

    def _encode(self, fn, blocks):
        return self.pool.map(fn, blocks) if self.pool is not None else map(fn, blocks)

    def write_patch(self, tile_y: int, tile_x: int, blocks, *, state: str = "valid") -> list[str]:
        if self.closed:
            raise ValueError("the batch writer is closed")
        key = self.product.key(tile_y, tile_x)
        if any(e["patch"] == key for e in self.pending):
            raise ValueError(f"patch {key} is already written")
        self.pending.append(self.product._store_patch(tile_y, tile_x, blocks, state, self._encode))
        return self.checkpoint() if len(self.pending) >= self.limit else []

    def write_invalid_patch(self, tile_y: int, tile_x: int, height: int, width: int,
                            fill) -> list[str]:
        """Every band holds ``fill``: one block encoded once for all bands."""
        writer = self.product.writer
        shape = (height, width, repr(fill))  # NaN fills compare equal by their repr
        payload = self.invalid_payloads.get(shape)
        if payload is None:
            payload = writer.encode_block(np.full((height, width), fill, dtype=writer.dtype))
            self.invalid_payloads[shape] = payload
        return self._write_encoded_patch(tile_y, tile_x, [payload] * writer.bands)

    def _write_encoded_patch(self, tile_y: int, tile_x: int, payloads: list[bytes]) -> list[str]:
        product, writer = self.product, self.product.writer
        key = product.key(tile_y, tile_x)
        if key in product.completed or any(e["patch"] == key for e in self.pending):
            raise ValueError(f"patch {key} is already written")
        by, bx = tile_y // PATCH_SIZE, tile_x // PATCH_SIZE
        written = []
        for band, payload in enumerate(payloads):
            index = writer.block_index(band, by, bx)
            writer.write_encoded(band, by, bx, payload)
            written.append([index, int(writer.offsets[index]), int(writer.counts[index])])
        self.pending.append({"patch": key, "state": "invalid", "blocks": written})
        return self.checkpoint() if len(self.pending) >= self.limit else []

    def checkpoint(self) -> list[str]:
        entries, self.pending = self.pending, []
        return self.product._checkpoint(entries)

    def close(self) -> list[str]:
        if self.closed:
            return []
        self.closed = True
        try:
            return self.checkpoint()
        finally:
            if self.pool is not None:
                self.pool.shutdown(wait=True)

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc, tb):
        self.close()
        return False
