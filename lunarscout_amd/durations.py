"""Mission-duration and safe-haven reductions of the lighting time series.

Drop-ins for the reference's landed-mission duration and safe-haven products
(paths under ``src/lunarscout/_numba_horizon/``):

* interval and sample helpers with the reference's rules and errors:
  ``CandidateStartInterval``, ``candidate_start_intervals``,
  ``monthly_candidate_intervals``, ``weekly_candidate_intervals``,
  ``fixed_candidate_intervals``, ``validate_evaluation_samples``
  (mission_duration.py:17-176) and ``build_month_bands``,
  ``month_indices_map`` (safe_haven.py:11-77);
* the reducers ``reduce_longest_candidate_duration_stream``
  (mission_duration.py:178-256) and ``reduce_safe_haven_patch_stream``
  (safe_haven.py:79-243) run fused into the per-pixel lighting walk
  (``lsh_duration_tiles``): the signal of each time is reduced as it is
  computed, with no [time][pixel] stream in memory. A ``DurationPlan`` holds the
  per-time tables both backends read;
* the body-elevation products ``run_sun_elevation_product`` /
  ``run_earth_elevation_product`` (elevation_pipeline.py:44-236): one float32
  band per time of the body centre's terrain-relative elevation;
* the products ``run_sunlight_duration_product``,
  ``run_sun_elevation_duration_product``,
  ``run_sunlight_earth_elevation_duration_product``,
  ``run_sun_elevation_earth_elevation_duration_product``
  (mission_duration_pipeline.py:335-405) and ``run_safe_haven_product``
  (safe_haven_pipeline.py:49-305): float32 GeoTIFF bands (one per candidate
  interval / month), NaN nodata, patches without a horizon tile left nodata.
  Explicit Moon-ME vectors replace the reference's SPICE lookups; the resumable
  product store (manifest, journal) is not part of this build.
"""

from __future__ import annotations

import bisect
import re
from collections.abc import Callable
from dataclasses import dataclass, field
from datetime import datetime, timedelta, timezone
from pathlib import Path
from typing import Literal

import numpy as np

from . import _lib
from .errors import (HipBackendError, HipError, HipExecutionError, InputError,
                     OperationCancelledError, ProductGenerationError)
from .geometry import DemGrid

DurationUnit = Literal["hours", "days"]


_ZULU = re.compile(r"Z$")


def _utc(value: datetime | str) -> datetime:
    """A timestamp as an aware UTC datetime (the rule of mission_duration.py:17-28).

    Accepts an aware datetime or an ISO-8601 string carrying its offset (a
    trailing ``Z`` is UTC); a naive timestamp is rejected.
    """
    stamp = (value if isinstance(value, datetime)
             else datetime.fromisoformat(_ZULU.sub("+00:00", str(value).strip())))
    offset = stamp.utcoffset()
    if offset is None:
        raise ValueError("mission-duration timestamps must include a UTC offset")
    return (stamp - offset).replace(tzinfo=timezone.utc)


@dataclass(frozen=True)
class CandidateStartInterval:
    """Half-open interval in which a qualifying mission may start (mission_duration.py:31-45)."""

    start_utc: datetime
    stop_utc: datetime

    def __post_init__(self) -> None:
        start, stop = _utc(self.start_utc), _utc(self.stop_utc)
        if start >= stop:
            raise ValueError("candidate-start interval must have positive duration")
        object.__setattr__(self, "start_utc", start)
        object.__setattr__(self, "stop_utc", stop)


def candidate_start_intervals(values, *, evaluation_start_utc, evaluation_stop_utc
                              ) -> tuple[CandidateStartInterval, ...]:
    """Normalise explicit intervals inside the evaluation range (mission_duration.py:48-75)."""
    start, stop = _utc(evaluation_start_utc), _utc(evaluation_stop_utc)
    if start >= stop:
        raise ValueError("evaluation interval must have positive duration")
    intervals = tuple(v if isinstance(v, CandidateStartInterval)
                      else CandidateStartInterval(_utc(v[0]), _utc(v[1])) for v in values)
    if not intervals:
        raise ValueError("at least one candidate-start interval is required")
    for interval in intervals:
        if interval.start_utc < start or interval.stop_utc > stop:
            raise ValueError("candidate-start intervals must be inside the evaluation interval")
    return intervals


def monthly_candidate_intervals(start_utc, stop_utc) -> tuple[CandidateStartInterval, ...]:
    """UTC calendar months clipped to ``[start, stop)`` (mission_duration.py:78-101)."""
    start, stop = _utc(start_utc), _utc(stop_utc)
    if start >= stop:
        raise ValueError("interval range must have positive duration")
    out = []
    year, month = start.year, start.month
    boundary = datetime(year, month, 1, tzinfo=timezone.utc)
    while boundary < stop:
        year, month = (year + 1, 1) if month == 12 else (year, month + 1)
        following = datetime(year, month, 1, tzinfo=timezone.utc)
        lo, hi = max(start, boundary), min(stop, following)
        if lo < hi:
            out.append(CandidateStartInterval(lo, hi))
        boundary = following
    return tuple(out)


def fixed_candidate_intervals(start_utc, stop_utc, *, duration: timedelta
                              ) -> tuple[CandidateStartInterval, ...]:
    """Fixed-length intervals anchored at ``start`` and clipped (mission_duration.py:111-131)."""
    start, stop = _utc(start_utc), _utc(stop_utc)
    if start >= stop:
        raise ValueError("interval range must have positive duration")
    if duration <= timedelta(0):
        raise ValueError("fixed interval duration must be positive")
    out, cursor = [], start
    while cursor < stop:
        following = min(stop, cursor + duration)
        out.append(CandidateStartInterval(cursor, following))
        cursor = following
    return tuple(out)


def weekly_candidate_intervals(start_utc, stop_utc) -> tuple[CandidateStartInterval, ...]:
    """Seven-day intervals anchored at ``start`` (mission_duration.py:104-108)."""
    return fixed_candidate_intervals(start_utc, stop_utc, duration=timedelta(days=7))


def validate_evaluation_samples(times_utc, *, evaluation_start_utc, evaluation_stop_utc
                                ) -> tuple[tuple[datetime, ...], np.ndarray]:
    """UTC samples and the hours each stands for, clipped at the stop
    (mission_duration.py:134-175): sample i covers [time[i], time[i + 1])."""
    start, stop = _utc(evaluation_start_utc), _utc(evaluation_stop_utc)
    if start >= stop:
        raise ValueError("evaluation interval must have positive duration")
    times = tuple(_utc(v) for v in times_utc)
    if not times:
        raise ValueError("at least one condition sample is required")
    if times[0] != start or times[-1] > stop:
        raise ValueError("samples must start at evaluation_start and not exceed evaluation_stop")
    if any(right <= left for left, right in zip(times, times[1:])):
        raise ValueError("sample timestamps must be strictly increasing")
    hours = np.array([((min(times[i + 1], stop) if i + 1 < len(times) else stop) - t)
                      .total_seconds() / 3600.0 for i, t in enumerate(times)], dtype=np.float64)
    return times, hours


def build_month_bands(times_utc) -> tuple[tuple[datetime, datetime], ...]:
    """Calendar-month ``[start, stop)`` bands covering the samples (safe_haven.py:22-50)."""
    stamps = tuple(_utc(v) for v in times_utc)
    if not stamps:
        return ()
    bands = []
    for year, month in sorted({(t.year, t.month) for t in stamps}):
        start = datetime(year, month, 1, tzinfo=timezone.utc)
        stop = datetime(year + 1, 1, 1, tzinfo=timezone.utc) if month == 12 else \
            datetime(year, month + 1, 1, tzinfo=timezone.utc)
        bands.append((start, stop))
    return tuple(bands)


def month_indices_map(times_utc, month_bands) -> np.ndarray:
    """int32 band index of every sample, -1 outside all bands (safe_haven.py:53-76)."""
    out = np.full(len(times_utc), -1, dtype=np.int32)
    for i, value in enumerate(times_utc):
        t = _utc(value)
        for b, (start, stop) in enumerate(month_bands):
            if start <= t < stop:
                out[i] = b
                break
    return out


@dataclass
class DurationPlan:
    """The per-time tables and thresholds of one reduction (``lsh_duration_spec``).

    ``interval_first[k]`` / ``interval_end[k]`` bound the samples whose time lies
    in candidate interval k; ``month_of[t]`` is the month band of sample t.
    """

    reducer: int
    time_count: int
    sun_threshold: float
    sun_signal: int = _lib.SIGNAL_SUN_FRACTION
    use_earth: bool = False
    earth_threshold_deg: float = 0.0
    sample_hours: np.ndarray | None = None
    interval_first: np.ndarray | None = None
    interval_end: np.ndarray | None = None
    output_days: bool = False
    month_of: np.ndarray | None = None
    band_count: int = 0
    step_hours: float = 0.0
    descriptions: list[str] = field(default_factory=list)

    @property
    def bands(self) -> int:
        return self.band_count if self.reducer == _lib.REDUCE_SAFE_HAVEN else len(self.interval_first)

    @property
    def needs_earth(self) -> bool:
        return self.reducer == _lib.REDUCE_SAFE_HAVEN or self.use_earth

    def identity(self) -> dict:
        """JSON identity of the plan (the resumable store's job configuration)."""
        def digest(a):
            return None if a is None else _sha256(a)

        return {"reducer": self.reducer, "time_count": self.time_count,
                "sun_signal": self.sun_signal, "sun_threshold": float(self.sun_threshold),
                "use_earth": bool(self.use_earth),
                "earth_threshold_deg": float(self.earth_threshold_deg),
                "sample_hours": digest(self.sample_hours),
                "interval_first": digest(self.interval_first),
                "interval_end": digest(self.interval_end), "output_days": bool(self.output_days),
                "month_of": digest(self.month_of), "band_count": int(self.band_count),
                "step_hours": float(self.step_hours)}

    def spec(self) -> _lib.DurationSpec:
        """The C struct; its pointers stay valid while this plan lives."""

        def ptr(a):
            return None if a is None else a.ctypes.data

        return _lib.DurationSpec(
            self.reducer, self.sun_signal, float(self.sun_threshold), int(self.use_earth),
            float(self.earth_threshold_deg), self.time_count, ptr(self.sample_hours),
            0 if self.interval_first is None else len(self.interval_first),
            ptr(self.interval_first), ptr(self.interval_end), int(self.output_days),
            ptr(self.month_of), int(self.band_count), float(self.step_hours))


def longest_duration_plan(times_utc, *, evaluation_start_utc, evaluation_stop_utc,
                          start_intervals, sun_signal: Literal["fraction", "margin"] = "fraction",
                          sun_threshold: float, earth_threshold_deg: float | None = None,
                          output_unit: DurationUnit = "hours") -> DurationPlan:
    """Validated plan of ``reduce_longest_candidate_duration_stream`` for these
    samples (mission_duration.py:178-215, mission_duration_pipeline.py:128-160)."""
    times, hours = validate_evaluation_samples(times_utc, evaluation_start_utc=evaluation_start_utc,
                                               evaluation_stop_utc=evaluation_stop_utc)
    intervals = candidate_start_intervals(start_intervals, evaluation_start_utc=evaluation_start_utc,
                                          evaluation_stop_utc=evaluation_stop_utc)
    if output_unit not in ("hours", "days"):
        raise ValueError("output_unit must be 'hours' or 'days'")
    if sun_signal not in ("fraction", "margin"):
        raise ValueError("sun_signal must be 'fraction' or 'margin'")
    if not np.isfinite(sun_threshold):
        raise ValueError("Sun threshold must be finite")
    if sun_signal == "fraction" and not 0.0 <= sun_threshold <= 1.0:
        raise ValueError("sunlight-fraction threshold must be between zero and one")
    if earth_threshold_deg is not None and not np.isfinite(earth_threshold_deg):
        raise ValueError("Earth elevation threshold must be finite")
    first = np.array([bisect.bisect_left(times, i.start_utc) for i in intervals], dtype=np.int32)
    end = np.array([bisect.bisect_left(times, i.stop_utc) for i in intervals], dtype=np.int32)
    return DurationPlan(
        reducer=_lib.REDUCE_LONGEST_DURATION, time_count=len(times),
        sun_threshold=float(sun_threshold),
        sun_signal=_lib.SIGNAL_SUN_FRACTION if sun_signal == "fraction" else _lib.SIGNAL_SUN_MARGIN,
        use_earth=earth_threshold_deg is not None,
        earth_threshold_deg=0.0 if earth_threshold_deg is None else float(earth_threshold_deg),
        sample_hours=np.ascontiguousarray(hours), interval_first=first, interval_end=end,
        output_days=output_unit == "days",
        descriptions=[f"{i.start_utc.isoformat()}/{i.stop_utc.isoformat()}" for i in intervals])


def safe_haven_plan(times_utc, *, sunlight_threshold: float, earth_threshold_deg: float,
                    time_step_hours: float, month_bands=None, month_index_of=None) -> DurationPlan:
    """Validated plan of ``reduce_safe_haven_patch_stream`` (safe_haven.py:79-158)."""
    if not times_utc:
        raise ValueError("time_count must be positive")
    if not 0.0 <= sunlight_threshold <= 1.0:
        raise ValueError("sunlight_threshold must be between zero and one")
    if not np.isfinite(earth_threshold_deg):
        raise ValueError("earth_threshold_deg must be finite")
    if not np.isfinite(time_step_hours) or time_step_hours <= 0.0:
        raise ValueError("time_step_hours must be positive and finite")
    bands = build_month_bands(times_utc) if month_bands is None else tuple(month_bands)
    months = (month_indices_map(times_utc, bands) if month_index_of is None
              else np.ascontiguousarray(month_index_of, dtype=np.int32))
    if months.shape != (len(times_utc),):
        raise ValueError("month_index_of must have shape (time_count,)")
    if len(bands) < 1:
        raise ValueError("at least one month band is required")
    if np.any((months < -1) | (months >= len(bands))):
        raise ValueError("month_index_of entries must be -1 or a band index")
    return DurationPlan(
        reducer=_lib.REDUCE_SAFE_HAVEN, time_count=len(times_utc),
        sun_threshold=float(sunlight_threshold), earth_threshold_deg=float(earth_threshold_deg),
        month_of=months, band_count=len(bands), step_hours=float(time_step_hours),
        descriptions=[f"{a.isoformat()}/{b.isoformat()}" for a, b in bands])


def duration_patch(calculator, dem: DemGrid, horizons_deg, plan: DurationPlan, sun_vectors_m,
                   earth_vectors_m=None, *, tile_y: int, tile_x: int, valid_height: int = 128,
                   valid_width: int = 128) -> tuple[np.ndarray, ...]:
    """One patch's reduced bands, float32 [valid_height, valid_width] each, as the
    reference reducers return them; ``calculator`` is a Hip- or CpuLightingSession."""
    return calculator.duration_patch(dem, horizons_deg, plan, sun_vectors_m, earth_vectors_m,
                                     tile_y=tile_y, tile_x=tile_x, valid_height=valid_height,
                                     valid_width=valid_width)


# -- products -------------------------------------------------------------------


def _float_product(operation: str, *, dem, horizon_store, output_path, band_count: int,
                   descriptions, compute, observer_elevation_m, nodata, compress, overwrite,
                   backend, progress_callback, progress_event_callback, cancellation_requested,
                   verbose, device_id, validate=None, configuration=None,
                   start_fresh: bool = False, timestamps_utc=()) -> Path:
    """Float32 multi-band product over the DEM's patches: ``compute(calculator, dem,
    tile, patch)`` yields the band blocks of one patch (the band order of the
    file); patches without a horizon tile are nodata. ``validate()`` runs after
    the argument checks and before any backend is touched. Written through the
    resumable store: an interrupted run of the same job resumes at the first
    patch not yet journaled (``start_fresh`` discards a staged product)."""
    from .lighting import select_lighting_session
    from .pipeline import enumerate_patches
    from .products import _dem, _Progress, drive_patches
    from .tiles import HorizonTileStore

    if backend not in ("auto", "cpu", "hip"):
        raise InputError("backend must be 'auto', 'cpu', or 'hip'.", code="product_backend_invalid",
                         details={"backend": backend})
    dem = _dem(dem)
    store = (horizon_store if isinstance(horizon_store, HorizonTileStore)
             else HorizonTileStore(horizon_store))
    output = Path(output_path).expanduser().resolve()
    if output.suffix.lower() not in (".tif", ".tiff"):
        raise InputError("Product output must use a .tif or .tiff extension.",
                         code="product_output_extension_invalid", details={"path": str(output)})
    if output.exists() and not overwrite:
        raise ProductGenerationError("The product output already exists.",
                                     code="product_output_exists", details={"path": str(output)})
    if validate is not None:
        validate()
    configuration = configuration() if callable(configuration) else (configuration or {})
    try:
        calculator, selected = select_lighting_session(backend, device_id=device_id)
    except HipBackendError as exc:
        raise HipError(f"The HIP {operation} backend is unavailable.",
                       code=f"hip_{operation}_unavailable", details={"error": str(exc)}) from exc
    progress = _Progress(operation, output, selected, verbose, progress_callback,
                         progress_event_callback)
    patches = enumerate_patches(dem.width, dem.height)
    from .product_store import ResumableTiledProduct, horizon_inventory_identity

    product = None
    inventory = horizon_inventory_identity(store, patches, observer_elevation_m)

    def check_cancel():
        if cancellation_requested is not None and cancellation_requested():
            raise OperationCancelledError(f"{operation} generation was cancelled.",
                                          code=f"{operation}_cancelled",
                                          details={"path": str(output)})

    def bands(patch, tile):
        for block in compute(calculator, dem, tile, patch):
            block = np.asarray(block, dtype=np.float32)
            if not np.isnan(nodata):
                block = np.where(np.isnan(block), np.float32(nodata), block)
            yield block

    try:
        product = ResumableTiledProduct(
            output, width=dem.width, height=dem.height, bands=band_count, dtype=np.float32,
            affine_transform=dem.geo_transform, projection=dem.projection, nodata=nodata,
            compress=compress, descriptions=descriptions, algorithm=operation,
            configuration=dict(configuration, observer_elevation_m=float(observer_elevation_m)),
            timestamps_utc=timestamps_utc, invalid_value=nodata,
            horizon_inventory_identity=inventory, backend=selected,
            overwrite=overwrite, start_fresh=start_fresh)
        progress.completed = len(product.completed)
        progress("start", progress.completed, len(patches))
        # tile reads, the fused reductions and the writes overlap (drive_patches)
        drive_patches(product, patches, store, observer_elevation_m, compute=bands,
                      progress=progress, check_cancel=check_cancel, invalid_fill=nodata)
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
        raise ProductGenerationError(f"{operation} file access failed.",
                                     code=f"{operation}_file_access_failed",
                                     details={"path": str(output), "error": str(exc)}) from exc
    except Exception as exc:
        if progress.callback_error is exc:
            raise
        raise ProductGenerationError(f"{operation} calculation failed.",
                                     code=f"{operation}_calculation_failed",
                                     details={"error": str(exc)}) from exc
    finally:
        if product is not None:
            product.suspend()  # the staged product stays for a resume


def _sha256(array) -> str | None:
    import hashlib

    if array is None:
        return None
    return hashlib.sha256(np.ascontiguousarray(array).tobytes()).hexdigest()


def _vectors(vectors_m, count, name):
    from .lighting import validate_vectors

    if vectors_m is None:
        return None
    try:
        vectors = validate_vectors(vectors_m)
    except ValueError as exc:
        raise InputError(str(exc), code="product_vectors_invalid") from exc
    if len(vectors) != count:
        raise InputError(f"timestamps and {name} vectors must align",
                         code="product_vectors_invalid",
                         details={"times": count, "vectors": len(vectors)})
    return vectors


def _run_product(operation: str, plan: DurationPlan, *, sun_vectors_m, earth_vectors_m,
                 **kwargs) -> Path:
    state = {}

    def validate():
        state["sun"] = _vectors(sun_vectors_m, plan.time_count, "Sun")
        state["earth"] = _vectors(earth_vectors_m, plan.time_count, "Earth")
        if state["sun"] is None:
            raise InputError("Sun vectors are required", code="product_vectors_required")
        if plan.needs_earth and state["earth"] is None:
            raise InputError("Earth vectors are required", code="product_vectors_required")

    def compute(calculator, dem, tile, patch):
        return calculator.duration_patch(dem, tile, plan, state["sun"], state["earth"],
                                         tile_y=patch.tile_y, tile_x=patch.tile_x,
                                         valid_height=patch.height, valid_width=patch.width)

    def configuration():
        return {"plan": plan.identity(), "sun_vectors_sha256": _sha256(state["sun"]),
                "earth_vectors_sha256": _sha256(state["earth"])}

    return _float_product(operation, band_count=plan.bands, descriptions=plan.descriptions,
                          compute=compute, validate=validate, configuration=configuration,
                          **kwargs)


def _duration_product(operation, *, sun_signal, sun_threshold, earth_threshold_deg, dem,
                      horizon_store, output_path, times_utc, evaluation_start_utc, evaluation_stop_utc,
                      start_intervals, sun_vectors_m, earth_vectors_m=None,
                      output_unit: DurationUnit = "hours", observer_elevation_m: float = 0.0,
                      nodata: float = float("nan"), compress: bool = True, overwrite: bool = False,
                      backend: str = "auto", progress_callback: Callable | None = None,
                      progress_event_callback: Callable | None = None,
                      cancellation_requested: Callable | None = None, verbose: bool = False,
                      device_id: int = 0, start_fresh: bool = False) -> Path:
    if earth_vectors_m is not None and earth_threshold_deg is None:
        raise ValueError("Earth elevation threshold must be finite")
    if earth_vectors_m is None and earth_threshold_deg is not None:
        raise ValueError("Earth vectors are required with an Earth threshold")
    plan = longest_duration_plan(times_utc, evaluation_start_utc=evaluation_start_utc,
                                 evaluation_stop_utc=evaluation_stop_utc,
                                 start_intervals=start_intervals, sun_signal=sun_signal,
                                 sun_threshold=sun_threshold,
                                 earth_threshold_deg=earth_threshold_deg, output_unit=output_unit)
    return _run_product(operation, plan, dem=dem, horizon_store=horizon_store,
                        output_path=output_path, sun_vectors_m=sun_vectors_m,
                        earth_vectors_m=earth_vectors_m, observer_elevation_m=observer_elevation_m, nodata=nodata, compress=compress,
                        overwrite=overwrite, backend=backend, progress_callback=progress_callback,
                        progress_event_callback=progress_event_callback,
                        cancellation_requested=cancellation_requested, verbose=verbose,
                        device_id=device_id, start_fresh=start_fresh)


def run_sunlight_duration_product(*, sunlight_fraction_threshold: float, **kwargs) -> Path:
    """Longest continuous sunlight-fraction duration per candidate-start interval
    (mission_duration_pipeline.py:335-342)."""
    return _duration_product("mission_duration", sun_signal="fraction",
                             sun_threshold=sunlight_fraction_threshold, earth_threshold_deg=None,
                             **kwargs)


def run_sun_elevation_duration_product(*, sun_elevation_threshold_deg: float, **kwargs) -> Path:
    """Longest continuous Sun-centre local-horizon margin duration (:345-352)."""
    return _duration_product("mission_duration", sun_signal="margin",
                             sun_threshold=sun_elevation_threshold_deg, earth_threshold_deg=None,
                             **kwargs)


def run_sunlight_earth_elevation_duration_product(*, sunlight_fraction_threshold: float,
                                                  earth_elevation_threshold_deg: float,
                                                  earth_vectors_m, **kwargs) -> Path:
    """Longest duration with sunlight fraction and Earth margin both met (:355-372)."""
    return _duration_product("mission_duration", sun_signal="fraction",
                             sun_threshold=sunlight_fraction_threshold,
                             earth_threshold_deg=earth_elevation_threshold_deg,
                             earth_vectors_m=earth_vectors_m, **kwargs)


def run_sun_elevation_earth_elevation_duration_product(*, sun_elevation_threshold_deg: float,
                                                       earth_elevation_threshold_deg: float,
                                                       earth_vectors_m, **kwargs) -> Path:
    """Longest duration with Sun-centre and Earth-centre margins both met (:375-392)."""
    return _duration_product("mission_duration", sun_signal="margin",
                             sun_threshold=sun_elevation_threshold_deg,
                             earth_threshold_deg=earth_elevation_threshold_deg,
                             earth_vectors_m=earth_vectors_m, **kwargs)


def run_safe_haven_product(*, dem, horizon_store, output_path, times_utc, sun_vectors_m,
                           earth_vectors_m, time_step_hours: float,
                           earth_threshold_deg: float = 2.0, sunlight_threshold: float = 0.2,
                           observer_elevation_m: float = 0.0,
                           nodata: float = float("nan"), compress: bool = True,
                           overwrite: bool = False, backend: str = "auto",
                           progress_callback: Callable | None = None,
                           progress_event_callback: Callable | None = None,
                           cancellation_requested: Callable | None = None, verbose: bool = False,
                           device_id: int = 0, start_fresh: bool = False) -> Path:
    """Per-month safe-haven durations (safe_haven_pipeline.py:49-305): the longest
    low-Sun run overlapping an Earth outage in the month, NaN where ill-posed."""
    plan = safe_haven_plan(times_utc, sunlight_threshold=sunlight_threshold,
                           earth_threshold_deg=earth_threshold_deg,
                           time_step_hours=time_step_hours)
    return _run_product("safe_haven", plan, dem=dem, horizon_store=horizon_store,
                        output_path=output_path, sun_vectors_m=sun_vectors_m,
                        earth_vectors_m=earth_vectors_m, observer_elevation_m=observer_elevation_m, nodata=nodata, compress=compress,
                        overwrite=overwrite, backend=backend, progress_callback=progress_callback,
                        progress_event_callback=progress_event_callback,
                        cancellation_requested=cancellation_requested, verbose=verbose,
                        device_id=device_id, start_fresh=start_fresh)


def _body_elevation_product(body: str, vectors_m, *, dem, horizon_store, output_path, times_utc,
                            observer_elevation_m: float = 0.0, nodata: float = float("nan"),
                            compress: bool = True, overwrite: bool = False, backend: str = "auto",
                            progress_callback: Callable | None = None,
                            progress_event_callback: Callable | None = None,
                            cancellation_requested: Callable | None = None,
                            verbose: bool = False, device_id: int = 0,
                            start_fresh: bool = False) -> Path:
    """One float32 band per time: the terrain-relative elevation of the body's
    centre, el - lerp(horizon) (elevation_pipeline.py:44-222)."""
    times = tuple(times_utc)
    state = {}

    def validate():
        state["vectors"] = _vectors(vectors_m, len(times), body)
        if state["vectors"] is None or not len(times):
            raise InputError(f"{body} vectors and timestamps are required",
                             code="product_vectors_required")

    def compute(calculator, dem, tile, patch):
        return calculator.iter_patch_margin_tiles(dem, tile, state["vectors"],
                                                  tile_y=patch.tile_y, tile_x=patch.tile_x,
                                                  valid_height=patch.height,
                                                  valid_width=patch.width)

    return _float_product(f"{body}_elevation", dem=dem, horizon_store=horizon_store,
                          output_path=output_path, band_count=len(times),
                          descriptions=[str(t) for t in times], compute=compute,
                          validate=validate, observer_elevation_m=observer_elevation_m,
                          timestamps_utc=times,
                          nodata=nodata, compress=compress, overwrite=overwrite, backend=backend,
                          progress_callback=progress_callback,
                          progress_event_callback=progress_event_callback,
                          cancellation_requested=cancellation_requested, verbose=verbose,
                          device_id=device_id, start_fresh=start_fresh,
                          configuration=lambda: {"vectors_sha256": _sha256(state["vectors"])})


def run_sun_elevation_product(*, sun_vectors_m, **kwargs) -> Path:
    """Sun-centre local-horizon elevation per time (elevation_pipeline.py:225-229)."""
    return _body_elevation_product("sun", sun_vectors_m, **kwargs)


def run_earth_elevation_product(*, earth_vectors_m, **kwargs) -> Path:
    """Earth-centre local-horizon elevation per time (elevation_pipeline.py:232-236)."""
    return _body_elevation_product("earth", earth_vectors_m, **kwargs)


__all__ = [
    "CandidateStartInterval", "DurationPlan", "build_month_bands", "candidate_start_intervals",
    "duration_patch", "fixed_candidate_intervals", "longest_duration_plan", "month_indices_map",
    "monthly_candidate_intervals", "run_earth_elevation_product", "run_safe_haven_product",
    "run_sun_elevation_duration_product", "run_sun_elevation_product",
    "run_sun_elevation_earth_elevation_duration_product", "run_sunlight_duration_product",
    "run_sunlight_earth_elevation_duration_product", "safe_haven_plan",
    "validate_evaluation_samples", "weekly_candidate_intervals",
]
