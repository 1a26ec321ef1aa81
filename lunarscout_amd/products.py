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
  whose horizon tile is missing carry ``invalid_value``;
* progress events and cancellation between patches.
"""

from __future__ import annotations

import sys
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

    def __call__(self, stage, completed, total, tile_y=None, tile_x=None):
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


def _run(operation, dem_path, horizons_path, output_path, *, sun_vectors_m, times, backend,
         observer_height_m, invalid_value, compress, overwrite, verbose, progress_callback,
         progress_event_callback, cancellation_requested, device_id, bands_of, per_patch,
         start_fresh=False):
    _validate_common(backend, {"progress_callback": progress_callback,
                               "progress_event_callback": progress_event_callback,
                               "cancellation_requested": cancellation_requested}, invalid_value)
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

    def blocks(tile, patch):
        for _, block in per_patch(calculator, dem, tile, prepared, patch):
            check_cancel()
            yield block

    try:
        import hashlib

        product = ResumableTiledProduct(
            output, width=dem.width, height=dem.height, bands=bands, dtype=np.uint8,
            affine_transform=dem.geo_transform, projection=dem.projection, compress=compress,
            descriptions=([str(t) for t in times]
                          if times is not None and len(times) == bands else None),
            algorithm=operation,
            configuration={"vectors_sha256": hashlib.sha256(vectors.tobytes()).hexdigest(),
                           "observer_height_m": float(observer_height_m)},
            timestamps_utc=(times if times is not None and len(times) == bands else ()),
            invalid_value=int(invalid_value), horizon_inventory_identity=inventory,
            backend=selected, overwrite=overwrite, start_fresh=start_fresh)
        progress("start", len(product.completed), len(patches))
        for k, patch in enumerate(patches):
            if product.is_complete(patch.tile_y, patch.tile_x):
                continue
            check_cancel()
            tile = store.read(patch.tile_y, patch.tile_x, observer_height_m)
            if tile is None:
                product.write_invalid_patch(patch.tile_y, patch.tile_x, patch.height,
                                            patch.width, int(invalid_value))
                progress("invalid", k + 1, len(patches), patch.tile_y, patch.tile_x)
                continue
            # a ray that met no terrain (-inf in a .bin tile) sees the whole sky
            tile = np.where(np.isfinite(tile), tile, np.float32(-90.0))
            product.write_patch(patch.tile_y, patch.tile_x, blocks(tile, patch))
            progress("valid", k + 1, len(patches), patch.tile_y, patch.tile_x)
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
                      start_fresh: bool = False) -> Path:
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
                bands_of=bands_of, per_patch=per_patch, start_fresh=start_fresh)


def generate_psr(dem_path, horizons_path, output_path, *, sun_vectors_m, times=None,
                 backend: str = "auto", observer_height_m: float = 0.0, invalid_value: int = 0,
                 compress: bool = True, overwrite: bool = False, verbose: bool = False,
                 progress_callback: Callable | None = None,
                 progress_event_callback: Callable | None = None,
                 cancellation_requested: Callable | None = None, device_id: int = 0,
                 start_fresh: bool = False) -> Path:
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
                bands_of=bands_of, per_patch=per_patch, start_fresh=start_fresh)
