"""``generate_horizons``: the public horizon generator on the HIP backend.

Drop-in for the reference's ``lunarscout.generate_horizons``
(``src/lunarscout/horizon.py:241-440``): the same arguments, validation and
error codes, the same 128x128-tile ``.cbin`` / ``.bin`` layout with resume on
structurally complete tiles, staged atomic tile writes, the sibling
``.pyr.bin`` pyramid cache, monotonic progress fractions and ``ProgressEvent``
stages, cancellation at tile boundaries. The backend literal is ``"hip"``.

Per tile the device runs the FP64 segment generator and the fused all-pass
march (``HipSession.horizon_tile``); only the degrees come back to the host,
where the writer thread encodes and publishes the tile while the GPU marches
the next one.
"""

from __future__ import annotations

import math
import os
import sys
import time
import uuid
from dataclasses import dataclass
from datetime import datetime, timedelta
from numbers import Real
from pathlib import Path
from typing import Any, Callable, Sequence

import numpy as np

from .errors import (HipBackendError, HipError, HipExecutionError, HorizonGenerationError,
                     InputError, OperationCancelledError)
from .geometry import DemGrid
from .pipeline import (HorizonPipelineCancelled, enumerate_patches, run_bounded_pipeline)
from .tiles import AZIMUTH_COUNT, PATCH_SIZE, HorizonTileStore


@dataclass(frozen=True)
class ProgressEvent:
    """One immutable progress observation (``src/lunarscout/progress.py``)."""

    operation: str
    stage: str
    completed: int
    total: int
    fraction: float
    backend: str | None
    message: str
    tile_y: int | None = None
    tile_x: int | None = None
    path: Path | None = None


class _ProgressAdapter:
    """Pipeline events -> monotonic fractions and ``ProgressEvent`` (horizon.py:44-133)."""

    def __init__(self, output: Path, *, verbose: bool, progress_callback, progress_event_callback):
        self.output = output
        self.verbose = verbose
        self.progress_callback = progress_callback
        self.progress_event_callback = progress_event_callback
        self.last: float | None = None
        self.started: float | None = None
        self.callback_error: BaseException | None = None

    def _timing(self, event: ProgressEvent) -> str:
        if event.stage == "process_patches" and self.started is None:
            self.started = time.monotonic()
        if self.started is None or event.completed <= 0:
            return "    --.-- s/patch | ETA -------------------"
        per = (time.monotonic() - self.started) / event.completed
        eta = datetime.now().astimezone() + timedelta(seconds=per * max(0, event.total - event.completed))
        return f"{per:8.2f} s/patch | ETA {eta:%Y-%m-%d %H:%M:%S}"

    def __call__(self, item) -> None:
        stage, done, total = str(item.stage), int(item.processed_patches), int(item.total_patches)
        if stage == "prepare_patches":
            fraction = 0.1
        elif stage == "process_patches":
            fraction = 0.15 + 0.85 * (0.0 if total == 0 else done / total)
        elif stage == "complete":
            fraction = 1.0
        else:
            fraction = min(1.0, max(0.0, float(item.percent) / 100.0))
        if self.last is not None:
            fraction = max(self.last, fraction)
        tile_y = tile_x = None
        if item.file_name:
            parts = str(item.file_name).split("_")
            if len(parts) >= 3:
                try:
                    tile_y, tile_x = int(parts[1]), int(parts[2])
                except ValueError:
                    pass
        event = ProgressEvent("horizons", stage, done, total, fraction, "hip", str(item.message),
                              tile_y, tile_x, self.output)
        if self.verbose:
            if stage == "prepare_patches":
                print("horizons: using hip backend", flush=True)
            print(f"horizons: {stage:<16} {done:6d}/{total:<6d} | {self._timing(event)}",
                  file=sys.stdout, flush=True)
        for callback, value in ((self.progress_event_callback, event),
                                (self.progress_callback, fraction)):
            if callback is None or (callback is self.progress_callback and fraction == self.last):
                continue
            try:
                callback(value)
            except BaseException as exc:
                self.callback_error = exc
                raise
        self.last = fraction


# -- pyramid cache (pyramid.py:34-75) ------------------------------------------------------

def pyramid_cache_path(dem_path) -> Path:
    """The C#-compatible sibling cache path of a DEM (pyramid.py:51-53)."""
    return Path(dem_path).with_suffix(".pyr.bin")


def _refresh_pyramid_cache(session, dem: DemGrid, dem_path: Path) -> None:
    """Build the pyramid on the device and publish the little-endian float32 cache.

    The device build takes microseconds, so the cache is never trusted for the
    march; it is rewritten (atomically) whenever its bytes differ from the
    pyramid of the DEM actually loaded -- a content check, where the reference
    checks only the length.
    """
    mips = session.build_max_pyramid(dem).mips.astype("<f4")
    path = pyramid_cache_path(dem_path)
    try:
        if path.is_file() and path.stat().st_size == mips.nbytes and path.read_bytes() == mips.tobytes():
            return
    except OSError:
        pass
    tmp = path.with_name(f".{path.name}.{uuid.uuid4().hex}.tmp")
    try:
        with tmp.open("wb") as handle:
            handle.write(mips.tobytes())
            handle.flush()
            os.fsync(handle.fileno())
        os.replace(tmp, path)
    except OSError:
        tmp.unlink(missing_ok=True)  # a read-only DEM directory keeps working without a cache


def _load_dems(dem_paths) -> tuple[list[DemGrid], list[Path | None]]:
    from .geotiff import load_dem

    dems, paths = [], []
    for item in dem_paths:
        if isinstance(item, DemGrid):
            dems.append(item)
            paths.append(None)
        else:
            dems.append(load_dem(item))
            paths.append(Path(item).expanduser().resolve())
    return dems, paths


def _validate(output_directory, dem_paths, observer_height_m, flags, callbacks, cancellation):
    if isinstance(dem_paths, (str, bytes, Path)):
        raise InputError("dem_paths must be a sequence ordered primary DEM first.",
                         code="horizon_dem_paths_invalid")
    try:
        items = [d if isinstance(d, DemGrid) else Path(d).expanduser().resolve() for d in dem_paths]
    except (TypeError, ValueError, OSError) as exc:
        raise InputError("dem_paths must contain filesystem paths.",
                         code="horizon_dem_paths_invalid") from exc
    if not items:
        raise InputError("At least one DEM path is required.", code="horizon_dem_paths_empty")
    if len(items) > 4:
        raise InputError("At most four DEMs are supported.", code="horizon_dem_paths_invalid",
                         details={"count": len(items)})
    missing = tuple(str(p) for p in items if isinstance(p, Path) and not p.is_file())
    if missing:
        raise InputError("A horizon DEM path does not identify a file.",
                         code="horizon_dem_not_found", details={"paths": missing})
    output = Path(output_directory).expanduser().resolve()
    if output.exists() and not output.is_dir():
        raise InputError("Horizon output must be a directory path.",
                         code="horizon_output_is_file", details={"path": str(output)})
    if (isinstance(observer_height_m, bool) or not isinstance(observer_height_m, Real)
            or not math.isfinite(observer_height_m) or not 0.0 <= float(observer_height_m) < 100.0):
        raise InputError("observer_height_m must be finite and between 0.0 and 100.0 meters.",
                         code="horizon_observer_height_invalid",
                         details={"observer_height_m": observer_height_m})
    for name, value in flags.items():
        if not isinstance(value, bool):
            raise InputError(f"{name} must be a bool.", code="horizon_option_invalid",
                             details={"argument": name})
    for name, callback in callbacks.items():
        if callback is not None and not callable(callback):
            raise InputError(f"{name} must be callable or None.", code="horizon_callback_invalid",
                             details={"argument": name})
    if cancellation is not None and cancellation():
        raise OperationCancelledError("Horizon generation was cancelled.",
                                      code="horizon_generation_cancelled",
                                      details={"path": str(output)})
    return output, items


def generate_horizons(output_directory, dem_paths: Sequence[Any], *, observer_height_m: float = 0.0,
                      compress: bool = True, overwrite: bool = False, verbose: bool = False,
                      progress_callback: Callable[[float], None] | None = None,
                      progress_event_callback: Callable[[ProgressEvent], None] | None = None,
                      cancellation_requested: Callable[[], bool] | None = None,
                      device_id: int = 0) -> Path:
    """Generate 128-pixel horizon tiles on an MI355X (``horizon.py:241-440``).

    ``dem_paths`` is ordered primary DEM first (GeoTIFF paths, or ``DemGrid``
    objects); later DEMs extend the terrain and inherit the running horizon for
    their culling. Tiles hold float32 degrees, 1440 azimuths at 0.25 degrees,
    sample 0 north, clockwise. Structurally complete tiles are skipped unless
    ``overwrite``. There is no CPU fallback: ``HipError`` with code
    ``hip_horizon_unavailable`` is raised before any tile is touched when the
    library or a device is missing.
    """
    output, _items = _validate(
        output_directory, dem_paths, observer_height_m,
        {"compress": compress, "overwrite": overwrite, "verbose": verbose},
        {"progress_callback": progress_callback,
         "progress_event_callback": progress_event_callback,
         "cancellation_requested": cancellation_requested},
        cancellation_requested)
    dems, paths = _load_dems(dem_paths)
    primary = dems[0]
    if ((primary.width - 1) // PATCH_SIZE) * PATCH_SIZE > 99_999 or \
            ((primary.height - 1) // PATCH_SIZE) * PATCH_SIZE > 99_999:
        raise InputError("The primary DEM exceeds the horizon filename coordinate range.",
                         code="horizon_grid_too_large",
                         details={"width": primary.width, "height": primary.height})
    try:
        from .hip_backend import HipSession

        session = HipSession(device_id)
    except (HipBackendError, HipExecutionError) as exc:
        raise HipError("A compatible AMD HIP device is required for horizon generation.",
                       code="hip_horizon_unavailable", details={"error": str(exc)}) from exc
    adapter = _ProgressAdapter(output, verbose=verbose, progress_callback=progress_callback,
                               progress_event_callback=progress_event_callback)
    height = float(observer_height_m)
    try:
        for dem, path in zip(dems, paths):
            session.dem_id(dem)
            if path is not None:
                _refresh_pyramid_cache(session, dem, path)
        output.mkdir(parents=True, exist_ok=True)

        def processor_factory(_worker_id: int):
            def process(patch, _prepared):
                return session.horizon_tile(dems, tile_column=patch.tile_x, tile_row=patch.tile_y,
                                            azimuth_count=AZIMUTH_COUNT,
                                            observer_elevation_m=height, degrees=True)
            return process

        run_bounded_pipeline(
            enumerate_patches(primary.width, primary.height), store=HorizonTileStore(output),
            prepare_patch=lambda _patch: None, processor_factory=processor_factory,
            observer_elevation_m=height, compress=compress, skip_existing=not overwrite,
            prepared_queue_capacity=1, writer_queue_capacity=1, worker_count=1,
            progress_callback=adapter, cancellation_requested=cancellation_requested)
        return output
    except Exception as exc:
        if adapter.callback_error is exc:
            raise
        if isinstance(exc, HorizonPipelineCancelled):
            raise OperationCancelledError("Horizon generation was cancelled.",
                                          code="horizon_generation_cancelled",
                                          details={"path": str(output)}) from exc
        if isinstance(exc, HipBackendError):
            raise HipError("A compatible AMD HIP device is required for horizon generation.",
                           code="hip_horizon_unavailable", details={"error": str(exc)}) from exc
        if isinstance(exc, HipExecutionError):
            raise HipError("HIP horizon generation failed during execution.",
                           code="hip_horizon_execution_failed", details={"error": str(exc)}) from exc
        raise HorizonGenerationError(
            "Horizon generation failed.",
            code="horizon_storage_failed" if isinstance(exc, OSError) else "horizon_generation_failed",
            details={"path": str(output), "error": str(exc)}) from exc
    finally:
        session.close()
