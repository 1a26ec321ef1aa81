"""Bounded tile pipeline: prepare -> compute -> finalize/write, with resume and cancel.

The reference's generator pipeline (``pipeline.py:88-450``) enumerates
row-major 128-pixel patches with partial right/bottom edges, skips tiles whose
file is already structurally complete, and streams the rest through a
preparation stage, GPU workers and a writer, each connected by a bounded queue,
checking for cancellation between units of work. This module keeps that
contract (``enumerate_patches``, ``run_bounded_pipeline`` with the same keyword
arguments, ``HorizonProgress`` events with the same stages) on a small generic
stage runner: every stage is a set of threads reading one bounded queue and
feeding the next, and the first error or a cancellation stops all of them.

With the HIP backend the preparation stage is empty (segments are generated on
the device inside the compute call), so the pipeline overlaps the GPU march of
tile k+1 with the encode and write of tile k.
"""

from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Callable, Sequence

import numpy as np

from .tiles import PATCH_SIZE, HorizonTileStore


class HorizonPipelineCancelled(RuntimeError):
    """Cancellation observed between bounded units of work (pipeline.py:23-24)."""


@dataclass(frozen=True)
class PatchDescriptor:
    """One 128x128 tile of the primary DEM (pipeline.py:27-45)."""

    index: int
    tile_x: int
    tile_y: int
    patch_x: int
    patch_y: int
    width: int
    height: int

    @property
    def kernel_width(self) -> int:
        return PATCH_SIZE

    @property
    def kernel_height(self) -> int:
        return PATCH_SIZE


@dataclass(frozen=True)
class HorizonProgress:
    processed_patches: int
    total_patches: int
    percent: float
    stage: str
    message: str
    file_name: str | None = None


@dataclass(frozen=True)
class PipelineResult:
    output_paths: tuple[Path, ...]
    skipped_patches: int
    wall_seconds: float
    preparation_seconds: float
    compute_seconds: float
    finalization_seconds: float
    write_seconds: float


def enumerate_patches(width: int, height: int, *, include_partial_edges: bool = True
                      ) -> list[PatchDescriptor]:
    """Row-major 128-pixel patches, partial right/bottom edges kept (pipeline.py:88-124)."""
    for v in (width, height):
        if isinstance(v, bool) or not isinstance(v, (int, np.integer)) or v <= 0:
            raise ValueError("DEM dimensions must be positive integers")
    if not include_partial_edges and (width % PATCH_SIZE or height % PATCH_SIZE):
        raise ValueError("DEM dimensions must be even multiples of 128")
    out = []
    for py in range((height + PATCH_SIZE - 1) // PATCH_SIZE):
        for px in range((width + PATCH_SIZE - 1) // PATCH_SIZE):
            x, y = px * PATCH_SIZE, py * PATCH_SIZE
            out.append(PatchDescriptor(len(out), x, y, px, py, min(PATCH_SIZE, width - x),
                                       min(PATCH_SIZE, height - y)))
    return out


_DONE = object()


@dataclass
class _Run:
    """Shared state of one pipeline run: the stop flag, the first error, timings."""

    cancelled: Callable[[], bool]
    stop: threading.Event = field(default_factory=threading.Event)
    error: BaseException | None = None
    lock: threading.Lock = field(default_factory=threading.Lock)
    seconds: dict = field(default_factory=lambda: {"prepare": 0.0, "compute": 0.0,
                                                   "finalize": 0.0, "write": 0.0})

    def fail(self, error: BaseException) -> None:
        with self.lock:
            if self.error is None:
                self.error = error
        self.stop.set()

    def check(self) -> None:
        if self.cancelled():
            raise HorizonPipelineCancelled("Horizon generation was cancelled.")

    def put(self, q: queue.Queue, item) -> bool:
        """Bounded put that gives up when the run stops (so no stage blocks forever)."""
        while True:
            try:
                q.put(item, timeout=0.05)
                return True
            except queue.Full:
                if self.stop.is_set():
                    return False

    def timed(self, name: str, fn, *args):
        started = time.perf_counter()
        result = fn(*args)
        with self.lock:
            self.seconds[name] += time.perf_counter() - started
        return result


def _stage(run: _Run, inbox: queue.Queue, outbox: queue.Queue | None, work, producers: int):
    """Thread body: apply ``work`` to items until every producer sent its end marker."""
    remaining = producers
    while remaining:
        item = inbox.get()
        if item is _DONE:
            remaining -= 1
            continue
        if run.stop.is_set():
            continue  # drain so upstream puts complete
        try:
            run.check()
            result = work(item)
            if outbox is not None and not run.put(outbox, result):
                continue
        except BaseException as error:  # noqa: BLE001 - the first error ends the run
            run.fail(error)


def run_bounded_pipeline(
    patches: Sequence[PatchDescriptor], *, store: HorizonTileStore,
    prepare_patch: Callable[[PatchDescriptor], Any],
    processor_factory: Callable[[int], Callable[[PatchDescriptor, Any], Any]],
    finalize_patch: Callable[[PatchDescriptor, Any], np.ndarray] | None = None,
    observer_elevation_m: float = 0.0, compress: bool = False, skip_existing: bool = True,
    prepared_queue_capacity: int = 1, writer_queue_capacity: int | None = None,
    worker_count: int = 1, progress_callback: Callable[[HorizonProgress], None] | None = None,
    cancellation_requested: Callable[[], bool] | None = None, progress_stream=None,
) -> PipelineResult:
    """Prepare, compute and write tiles with bounded memory (pipeline.py:127-450)."""
    if prepared_queue_capacity < 1:
        raise ValueError("prepared_queue_capacity must be positive")
    if worker_count < 1:
        raise ValueError("worker_count must be positive")
    if writer_queue_capacity is not None and writer_queue_capacity < 1:
        raise ValueError("writer_queue_capacity must be positive or None")
    started = time.perf_counter()
    run = _Run(cancellation_requested or (lambda: False))
    run.check()
    finalize_patch = finalize_patch or (lambda _patch, payload: payload)
    pending = [p for p in patches
               if not (skip_existing and store.find_existing_path(p.tile_y, p.tile_x,
                                                                  observer_elevation_m))]
    skipped = len(patches) - len(pending)
    total = len(pending)
    written: list[Path] = []

    def emit(event: HorizonProgress) -> None:
        if progress_callback is not None:
            progress_callback(event)
        if progress_stream is not None:
            progress_stream.write(f"[{event.stage}] {event.percent:5.1f}% {event.message}\n")
            progress_stream.flush()

    emit(HorizonProgress(0, total, 10.0, "prepare_patches", "Preparing horizon patch pipeline."))
    if total == 0:
        emit(HorizonProgress(0, 0, 100.0, "complete", "No horizon patches need to be generated."))
        return PipelineResult((), skipped, time.perf_counter() - started, 0.0, 0.0, 0.0, 0.0)
    emit(HorizonProgress(0, total, 15.0, "process_patches", "Starting horizon patch generation."))

    prepared: queue.Queue = queue.Queue(maxsize=prepared_queue_capacity)
    computed: queue.Queue = queue.Queue(maxsize=writer_queue_capacity or 1)

    def write(item) -> None:
        patch, payload = item
        run.check()
        degrees = run.timed("finalize", finalize_patch, patch, payload)
        run.check()
        path = run.timed("write", lambda: store.write(
            patch.tile_y, patch.tile_x, observer_elevation_m, degrees, compress=compress,
            valid_width=patch.width, valid_height=patch.height))
        with run.lock:
            written.append(path)
            done = len(written)
        emit(HorizonProgress(done, total, done * 100.0 / total, "process_patches",
                             f"Generated {done}/{total} horizon patches.", path.name))

    def producer() -> None:
        try:
            for patch in pending:
                if run.stop.is_set():
                    break
                run.check()
                payload = run.timed("prepare", prepare_patch, patch)
                if not run.put(prepared, (patch, payload)):
                    break
        except BaseException as error:  # noqa: BLE001
            run.fail(error)
        finally:
            for _ in range(worker_count):
                prepared.put(_DONE)

    def worker(worker_id: int) -> None:
        try:
            processor = processor_factory(worker_id)
        except BaseException as error:  # noqa: BLE001
            run.fail(error)
            processor = None

        def compute(item):
            if processor is None:
                raise RuntimeError("processor unavailable")
            patch, payload = item
            result = run.timed("compute", processor, patch, payload)
            run.check()
            if writer_queue_capacity is None:  # finalize and write on the worker
                write((patch, result))
                return None
            return patch, result

        _stage(run, prepared, computed if writer_queue_capacity is not None else None, compute, 1)
        if writer_queue_capacity is not None:
            computed.put(_DONE)

    threads = [threading.Thread(target=producer, name="horizon-prepare")]
    threads += [threading.Thread(target=worker, args=(i,), name=f"horizon-worker-{i}")
                for i in range(worker_count)]
    if writer_queue_capacity is not None:
        threads.append(threading.Thread(target=_stage, name="horizon-writer",
                                        args=(run, computed, None, write, worker_count)))
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if run.error is not None:
        raise run.error
    emit(HorizonProgress(total, total, 100.0, "complete", "Horizons generation complete."))
    s = run.seconds
    return PipelineResult(tuple(written), skipped, time.perf_counter() - started, s["prepare"],
                          s["compute"], s["finalize"], s["write"])
