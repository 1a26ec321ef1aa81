"""Lightmap and PSR products from stored horizon tiles, ``backend="cpu"|"hip"|"auto"``.

Drop-ins for the reference's ``generate_lightmap`` and ``generate_psr``
(``src/lunarscout/products.py:315-698``, pipelines ``lightmap_pipeline.py`` and
``psr_pipeline.py``) on this path:

* inputs: the DEM (GeoTIFF path or ``DemGrid``), the horizon tile directory
  (``.cbin`` / ``.bin`` tiles as ``generate_horizons`` writes them), explicit
  Moon-ME Sun vectors (the reference can also derive them from SPICE, which is
  outside this build);
* backend selection as the reference: ``"cpu"`` never probes the GPU, ``"hip"``
  never falls back, ``"auto"`` uses the GPU when a session can be created;
* output: a tiled 128x128 deflate GeoTIFF (BigTIFF when large) written through
  a staged file and an atomic rename: the lightmap holds one uint8 band per
  time, trunc(255 * visible fraction) of the 16-slice 0.27-degree Sun disk; the
  PSR mask one uint8 band (255 = permanently shadowed). Pixels of patches
  whose horizon tile is missing or unreadable carry ``invalid_value``;
  ``output_transform`` / ``output_dtype`` / ``output_transform_id`` convert
  valid patches before they are stored (products.py:361-372);
* ``pipeline_mode="bounded"`` (the default) overlaps tile reads and decodes
  (``reader_worker_count`` threads, at most ``decoded_horizon_capacity``
  decoded tiles alive), the calculation on the calling thread and the writer
  thread (``writer_queue_capacity`` computed patches queued for it), which
  compresses band blocks on a thread pool and makes patches durable
  ``durable_batch_size`` at a time (psr_pipeline.py:126-160, 523-797);
  ``"serial"`` runs the same steps one patch after another;
* progress events -- start, read, calculate, write, then valid / invalid when
  a patch is durable, complete -- and cancellation between patches and bands.
"""

from __future__ import annotations

import os
import queue
import sys
import threading
from pathlib import Path
from typing import Callable

import numpy as np

from .errors import (HipBackendError, HipError, HipExecutionError, InputError,
                     OperationCancelledError, ProductGenerationError)
from .geometry import DemGrid
from .horizon import ProgressEvent
from .lighting import reduce_sun_vectors_for_psr, select_lighting_session, validate_vectors
from .pipeline import enumerate_patches
from .tiles import HorizonTileStore

BACKENDS = ("auto", "cpu", "hip")


def _preflight(horizons_path, output_path, overwrite: bool) -> tuple[Path, Path]:
    """products.py:180-212."""
    horizons = Path(horizons_path).expanduser().resolve()
    output = Path(output_path).expanduser().resolve()
    if not horizons.is_dir():
        raise InputError("The horizons path does not identify a directory.",
                         code="product_horizons_not_found", details={"path": str(horizons)})
    if output.suffix.lower() not in (".tif", ".tiff"):
        raise InputError("Product output must use a .tif or .tiff extension.",
                         code="product_output_extension_invalid", details={"path": str(output)})
    if output.exists() and not overwrite:
        raise ProductGenerationError("The product output already exists.",
                                     code="product_output_exists", details={"path": str(output)})
    if output.exists() and not output.is_file():
        raise ProductGenerationError("The product output path is not a file.",
                                     code="product_output_invalid", details={"path": str(output)})
    return horizons, output


def _validate_common(backend, callbacks, invalid_value):
    if backend not in BACKENDS:
        raise InputError("backend must be 'auto', 'cpu', or 'hip'.", code="product_backend_invalid",
                         details={"backend": backend})
    for name, cb in callbacks.items():
        if cb is not None and not callable(cb):
            raise InputError(f"{name} must be callable or None.", code="product_callback_invalid",
                             details={"argument": name})
    if isinstance(invalid_value, bool) or not isinstance(invalid_value, (int, np.integer)) or \
            not 0 <= int(invalid_value) <= 255:
        raise InputError("invalid_value must be an integer in [0, 255].",
                         code="product_invalid_value_invalid", details={"invalid_value": invalid_value})


def _vectors(sun_vectors_m, times):
    if sun_vectors_m is None:
        raise InputError("Explicit Moon-ME Sun vectors are required (no ephemeris in this build).",
                         code="product_vectors_required")
    try:
        vectors = validate_vectors(sun_vectors_m)
    except ValueError as exc:
        raise InputError(str(exc), code="product_vectors_invalid") from exc
    if times is not None and len(times) != len(vectors):
        raise InputError("The vector count must match the time count.",
                         code="product_vectors_invalid",
                         details={"times": len(times), "vectors": len(vectors)})
    return vectors


def _dem(dem_path) -> DemGrid:
    if isinstance(dem_path, DemGrid):
        return dem_path
    from .geotiff import load_dem

    return load_dem(dem_path)


class _Progress:
    def __init__(self, operation, output, backend, verbose, progress_callback, event_callback):
        self.operation, self.output, self.backend, self.verbose = operation, output, backend, verbose
        self.progress_callback, self.event_callback = progress_callback, event_callback
        self.last = None
        self.callback_error: BaseException | None = None
        self.completed = 0  # durable patches, advanced by the writer
        self.lock = threading.Lock()  # the pipeline's stages report from their threads

    def __call__(self, stage, completed, total, tile_y=None, tile_x=None):
        with self.lock:
            self._emit(stage, completed, total, tile_y, tile_x)

    def _emit(self, stage, completed, total, tile_y, tile_x):
        fraction = 1.0 if total == 0 else completed / total
        event = ProgressEvent(self.operation, stage, completed, total, fraction, self.backend,
                              f"{self.operation} {stage}", tile_y, tile_x, self.output)
        if self.verbose:
            if stage == "start":
                print(f"{self.operation}: using {self.backend} backend", file=sys.stdout, flush=True)
            elif stage in ("valid", "invalid", "complete"):
                print(f"{self.operation}: {stage} {completed}/{total}", file=sys.stdout, flush=True)
        try:
            if self.event_callback is not None:
                self.event_callback(event)
            if self.progress_callback is not None and completed != self.last:
                self.progress_callback(fraction)
                self.last = completed
        except BaseException as exc:
            self.callback_error = exc
            raise


def _validate_output_conversion(output_transform, output_dtype, output_transform_id):
    """products.py:36-70: the optional per-patch conversion of valid pixels."""
    if output_transform is None:
        if output_dtype is not None or output_transform_id is not None:
            raise InputError("output_dtype and output_transform_id require output_transform.",
                             code="product_output_conversion_invalid")
        return np.dtype(np.uint8)
    if not callable(output_transform):
        raise InputError("output_transform must be callable or None.",
                         code="product_output_conversion_invalid")
    if output_dtype is None:
        raise InputError("output_dtype is required with output_transform.",
                         code="product_output_conversion_invalid")
    try:
        dtype = np.dtype(output_dtype)
    except (TypeError, ValueError) as exc:
        raise InputError("output_dtype is not a valid NumPy dtype.",
                         code="product_output_dtype_invalid",
                         details={"output_dtype": repr(output_dtype)}) from exc
    if output_transform_id is not None and not isinstance(output_transform_id, str):
        raise InputError("output_transform_id must be a string or None.",
                         code="product_output_conversion_invalid")
    return dtype


def _validate_pipeline(mode, reader_workers, decoded_capacity, writer_capacity, batch):
    """The bounded pipeline's knobs (psr_pipeline.py:126-160)."""
    if mode not in ("serial", "bounded"):
        raise InputError("pipeline_mode must be 'serial' or 'bounded'.",
                         code="product_pipeline_invalid", details={"pipeline_mode": mode})
    for name, value in (("reader_worker_count", reader_workers),
                        ("decoded_horizon_capacity", decoded_capacity),
                        ("writer_queue_capacity", writer_capacity),
                        ("durable_batch_size", batch)):
        if isinstance(value, bool) or not isinstance(value, (int, np.integer)) or value < 1:
            raise InputError(f"{name} must be a positive integer.",
                             code="product_pipeline_invalid", details={name: value})


class _Stop(Exception):
    """Raised inside a stage when another stage has already failed."""


class _Overlap:
    """Read -> compute -> write over the pending patches, bounded.

    Horizon tiles are read and decoded on ``reader_workers`` threads (the codec
    runs in C and releases the GIL), at most ``decoded_capacity`` decoded tiles
    alive at once, handed over in patch order; the calling thread computes each
    patch (the GPU launches stay on one thread); one writer thread encodes and
    stores the band blocks and checkpoints them ``durable_batch_size`` at a time
    (psr_pipeline.py:523-797 runs its reader, CUDA and writer stages so). The
    first error of any stage stops the others and is raised by ``run``.
    """

    def __init__(self, pending, read, compute, write, finish, *, reader_workers, decoded_capacity,
                 writer_capacity):
        self.pending, self.read, self.compute = pending, read, compute
        self.write, self.finish = write, finish
        self.reader_workers = reader_workers
        self.slots = threading.BoundedSemaphore(decoded_capacity)
        self.decoded: queue.Queue = queue.Queue()
        self.computed: queue.Queue = queue.Queue(maxsize=writer_capacity)
        self.stop = threading.Event()
        self.errors: list[BaseException] = []
        self.lock = threading.Lock()

    def fail(self, error: BaseException) -> None:
        with self.lock:
            if not self.errors:
                self.errors.append(error)
        self.stop.set()

    def _put(self, q: queue.Queue, item) -> None:
        while True:
            try:
                q.put(item, timeout=0.05)
                return
            except queue.Full:
                if self.stop.is_set():
                    raise _Stop() from None

    def _get(self, q: queue.Queue):
        while True:
            try:
                return q.get(timeout=0.05)
            except queue.Empty:
                if self.stop.is_set():
                    raise _Stop() from None

    def _reader(self) -> None:
        from concurrent.futures import ThreadPoolExecutor

        try:
            with ThreadPoolExecutor(self.reader_workers, thread_name_prefix="product-read") as pool:
                window = []  # reads in flight, in patch order
                for patch in self.pending:
                    # a decoded-tile slot, released when the patch is computed; while
                    # none is free, hand the oldest read over so one can be
                    while not self.slots.acquire(timeout=0.05):
                        if self.stop.is_set():
                            raise _Stop()
                        if window:
                            self._put(self.decoded, window.pop(0).result())
                    window.append(pool.submit(self.read, patch))
                    while window and (window[0].done() or len(window) >= self.reader_workers):
                        self._put(self.decoded, window.pop(0).result())
                while window:
                    self._put(self.decoded, window.pop(0).result())
        except _Stop:
            pass
        except BaseException as error:  # noqa: BLE001 - the first error ends the run
            self.fail(error)
        finally:
            self.decoded.put(_END)

    def _writer(self) -> None:
        try:
            while True:
                item = self._get(self.computed)
                if item is _END:
                    break
                self.write(*item)
            self.finish()
        except _Stop:
            pass
        except BaseException as error:  # noqa: BLE001
            self.fail(error)

    def run(self) -> None:
        reader = threading.Thread(target=self._reader, name="product-reader")
        writer = threading.Thread(target=self._writer, name="product-writer")
        reader.start()
        writer.start()
        try:
            while True:
                item = self._get(self.decoded)
                if item is _END:
                    break
                patch, tile = item
                try:
                    result = self.compute(patch, tile)
                finally:
                    self.slots.release()
                self._put(self.computed, (patch, result))
            self._put(self.computed, _END)
        except _Stop:
            pass
        except BaseException as error:  # noqa: BLE001
            self.fail(error)
        finally:
            reader.join()
            writer.join()
        if self.errors:
            raise self.errors[0]


_END = object()


def drive_patches(product, patches, store, observer_height_m, *, compute, progress, check_cancel,
                  invalid_fill, pipeline_mode="bounded", reader_worker_count=4,
                  decoded_horizon_capacity=5, writer_queue_capacity=1,
                  durable_batch_size=16) -> None:
    """Every patch of ``patches`` not yet complete in ``product``: read its
    horizon tile, ``compute(patch, degrees)`` -> band blocks (None when the
    tile is missing or unreadable: the patch is written with ``invalid_fill``),
    write and checkpoint, serially or overlapped (``_Overlap``). ``progress``
    receives read / calculate / write events and, when a patch is durable, its
    state; ``progress.completed`` counts durable patches. Whole patches written
    before an error are checkpointed before it propagates."""
    pending = [p for p in patches if not product.is_complete(p.tile_y, p.tile_x)]
    by_key = {product.key(p.tile_y, p.tile_x): p for p in pending}
    total = len(patches)

    def read(patch):
        """(patch, degrees [pixels, N] or None when the tile is missing or unreadable)."""
        progress("read", progress.completed, total, patch.tile_y, patch.tile_x)
        try:
            tile = store.read(patch.tile_y, patch.tile_x, observer_height_m)
        except (OSError, ValueError):  # psr_pipeline.py:384-389: an unreadable tile is invalid
            tile = None
        if tile is not None:
            # a ray that met no terrain (-inf in a .bin tile) sees the whole sky
            tile = np.where(np.isfinite(tile), tile, np.float32(-90.0))
        return patch, tile

    def calculate(patch, tile):
        check_cancel()
        if tile is None:
            return None
        progress("calculate", progress.completed, total, patch.tile_y, patch.tile_x)
        blocks = []
        for block in compute(patch, tile):
            check_cancel()
            blocks.append(np.asarray(block))
        return blocks

    bands = product.writer.bands
    encode_workers = min(8, os.cpu_count() or 1) if bands > 1 else 0
    with product.batch_writer(int(durable_batch_size), encode_workers=encode_workers) as batch:

        def durable(keys):
            for key in keys:
                patch = by_key[key]
                progress.completed += 1
                progress(product.completed[key], progress.completed, total, patch.tile_y,
                         patch.tile_x)

        def write(patch, blocks):
            progress("write", progress.completed, total, patch.tile_y, patch.tile_x)
            if blocks is None:
                durable(batch.write_invalid_patch(patch.tile_y, patch.tile_x, patch.height,
                                                  patch.width, invalid_fill))
            else:
                durable(batch.write_patch(patch.tile_y, patch.tile_x, blocks))

        def finish():
            durable(batch.checkpoint())

        try:
            if pipeline_mode == "serial":
                for patch in pending:
                    check_cancel()
                    _, tile = read(patch)
                    write(patch, calculate(patch, tile))
                finish()
            else:
                _Overlap(pending, read, calculate, write, finish,
                         reader_workers=int(reader_worker_count),
                         decoded_capacity=int(decoded_horizon_capacity),
                         writer_capacity=int(writer_queue_capacity)).run()
        finally:
            durable(batch.checkpoint())  # whole patches written before an error stay


def _run(operation, dem_path, horizons_path, output_path, *, sun_vectors_m, times, backend,
         observer_height_m, invalid_value, compress, overwrite, verbose, progress_callback,
         progress_event_callback, cancellation_requested, device_id, bands_of, per_patch,
         start_fresh=False, output_transform=None, output_dtype=None, output_transform_id=None,
         pipeline_mode="bounded", reader_worker_count=4, decoded_horizon_capacity=5,
         writer_queue_capacity=1, durable_batch_size=16):
    _validate_common(backend, {"progress_callback": progress_callback,
                               "progress_event_callback": progress_event_callback,
                               "cancellation_requested": cancellation_requested}, invalid_value)
    dtype = _validate_output_conversion(output_transform, output_dtype, output_transform_id)
    _validate_pipeline(pipeline_mode, reader_worker_count, decoded_horizon_capacity,
                       writer_queue_capacity, durable_batch_size)
    horizons, output = _preflight(horizons_path, output_path, overwrite)
    dem = _dem(dem_path)
    vectors = _vectors(sun_vectors_m, times)
    try:
        calculator, selected = select_lighting_session(backend, device_id=device_id)
    except HipBackendError as exc:
        raise HipError(f"The HIP {operation} backend is unavailable.",
                       code=f"hip_{operation}_unavailable", details={"error": str(exc)}) from exc
    progress = _Progress(operation, output, selected, verbose, progress_callback,
                         progress_event_callback)
    store = HorizonTileStore(horizons)
    patches = enumerate_patches(dem.width, dem.height)
    from .product_store import ResumableTiledProduct, horizon_inventory_identity

    bands, prepared = bands_of(dem, vectors)
    inventory = horizon_inventory_identity(store, patches, observer_height_m)
    product = None

    def check_cancel():
        if cancellation_requested is not None and cancellation_requested():
            raise OperationCancelledError(f"{operation.capitalize()} generation was cancelled.",
                                          code=f"{operation}_cancelled",
                                          details={"path": str(output)})

    try:
        import hashlib

        configuration = {"vectors_sha256": hashlib.sha256(vectors.tobytes()).hexdigest(),
                         "observer_height_m": float(observer_height_m)}
        if output_transform is not None:
            configuration["output_transform_id"] = output_transform_id
        product = ResumableTiledProduct(
            output, width=dem.width, height=dem.height, bands=bands, dtype=dtype,
            affine_transform=dem.geo_transform, projection=dem.projection, compress=compress,
            descriptions=([str(t) for t in times]
                          if times is not None and len(times) == bands else None),
            algorithm=operation, configuration=configuration,
            timestamps_utc=(times if times is not None and len(times) == bands else ()),
            invalid_value=int(invalid_value), horizon_inventory_identity=inventory,
            backend=selected, overwrite=overwrite, start_fresh=start_fresh,
            output_transform=output_transform)
        progress.completed = len(product.completed)
        progress("start", progress.completed, len(patches))
        drive_patches(product, patches, store, observer_height_m,
                      compute=lambda patch, tile: (block for _, block in
                                                   per_patch(calculator, dem, tile, prepared,
                                                             patch)),
                      progress=progress, check_cancel=check_cancel,
                      invalid_fill=int(invalid_value), pipeline_mode=pipeline_mode,
                      reader_worker_count=reader_worker_count,
                      decoded_horizon_capacity=decoded_horizon_capacity,
                      writer_queue_capacity=writer_queue_capacity,
                      durable_batch_size=durable_batch_size)
        check_cancel()
        path = product.finalize()
        product = None
        progress("complete", len(patches), len(patches))
        return path
    except (OperationCancelledError, InputError, ProductGenerationError):
        raise
    except HipExecutionError as exc:
        raise HipError(f"The HIP {operation} backend failed during execution.",
                       code=f"hip_{operation}_execution_failed", details={"error": str(exc)}) from exc
    except OSError as exc:
        raise ProductGenerationError(f"{operation.capitalize()} file access failed.",
                                     code=f"{operation}_file_access_failed",
                                     details={"path": str(output), "error": str(exc)}) from exc
    except Exception as exc:
        if progress.callback_error is exc:
            raise
        raise ProductGenerationError(f"{operation.capitalize()} calculation failed.",
                                     code=f"{operation}_calculation_failed",
                                     details={"error": str(exc)}) from exc
    finally:
        if product is not None:
            product.suspend()  # the staged product stays for a resume


def generate_lightmap(dem_path, horizons_path, output_path, *, sun_vectors_m, times=None,
                      backend: str = "auto", observer_height_m: float = 0.0,
                      invalid_value: int = 0, compress: bool = True, overwrite: bool = False,
                      verbose: bool = False, progress_callback: Callable | None = None,
                      progress_event_callback: Callable | None = None,
                      cancellation_requested: Callable | None = None, device_id: int = 0,
                      start_fresh: bool = False,
                      output_transform: Callable | None = None, output_dtype=None,
                      output_transform_id: str | None = None,
                      pipeline_mode: str = "bounded", reader_worker_count: int = 4,
                      decoded_horizon_capacity: int = 5, writer_queue_capacity: int = 1,
                      durable_batch_size: int = 16) -> Path:
    """Tiled uint8 visible-solar-fraction GeoTIFF, one band per time (products.py:315-518)."""

    def bands_of(dem, vectors):
        return len(vectors), vectors

    def per_patch(calculator, dem, tile, vectors, patch):
        tiles = calculator.iter_patch_tiles(dem, tile, vectors, tile_y=patch.tile_y,
                                            tile_x=patch.tile_x, valid_height=patch.height,
                                            valid_width=patch.width)
        return enumerate(tiles)

    return _run("lightmap", dem_path, horizons_path, output_path, sun_vectors_m=sun_vectors_m,
                times=times, backend=backend, observer_height_m=observer_height_m,
                invalid_value=invalid_value, compress=compress, overwrite=overwrite,
                verbose=verbose, progress_callback=progress_callback,
                progress_event_callback=progress_event_callback,
                cancellation_requested=cancellation_requested, device_id=device_id,
                bands_of=bands_of, per_patch=per_patch, start_fresh=start_fresh,
                output_transform=output_transform, output_dtype=output_dtype,
                output_transform_id=output_transform_id, pipeline_mode=pipeline_mode,
                reader_worker_count=reader_worker_count,
                decoded_horizon_capacity=decoded_horizon_capacity,
                writer_queue_capacity=writer_queue_capacity,
                durable_batch_size=durable_batch_size)


def generate_psr(dem_path, horizons_path, output_path, *, sun_vectors_m, times=None,
                 backend: str = "auto", observer_height_m: float = 0.0, invalid_value: int = 0,
                 compress: bool = True, overwrite: bool = False, verbose: bool = False,
                 progress_callback: Callable | None = None,
                 progress_event_callback: Callable | None = None,
                 cancellation_requested: Callable | None = None, device_id: int = 0,
                 start_fresh: bool = False,
                 output_transform: Callable | None = None, output_dtype=None,
                 output_transform_id: str | None = None,
                 pipeline_mode: str = "bounded", reader_worker_count: int = 4,
                 decoded_horizon_capacity: int = 5, writer_queue_capacity: int = 1,
                 durable_batch_size: int = 16) -> Path:
    """Single-band uint8 PSR classification GeoTIFF, 255 = PSR (products.py:520-698).

    The Sun vectors are first reduced to the per-azimuth-bin maxima seen from
    the DEM's corners and centre (psr.py:61-91).
    """

    def bands_of(dem, vectors):
        reduced, _ = reduce_sun_vectors_for_psr(dem, vectors)
        return 1, reduced

    def per_patch(calculator, dem, tile, reduced, patch):
        mask = calculator.compute_patch(dem, tile, reduced, tile_y=patch.tile_y,
                                        tile_x=patch.tile_x, valid_height=patch.height,
                                        valid_width=patch.width)
        return [(0, mask)]

    return _run("psr", dem_path, horizons_path, output_path, sun_vectors_m=sun_vectors_m,
                times=times, backend=backend, observer_height_m=observer_height_m,
                invalid_value=invalid_value, compress=compress, overwrite=overwrite,
                verbose=verbose, progress_callback=progress_callback,
                progress_event_callback=progress_event_callback,
                cancellation_requested=cancellation_requested, device_id=device_id,
                bands_of=bands_of, per_patch=per_patch, start_fresh=start_fresh,
                output_transform=output_transform, output_dtype=output_dtype,
                output_transform_id=output_transform_id, pipeline_mode=pipeline_mode,
                reader_worker_count=reader_worker_count,
                decoded_horizon_capacity=decoded_horizon_capacity,
                writer_queue_capacity=writer_queue_capacity,
                durable_batch_size=durable_batch_size)
