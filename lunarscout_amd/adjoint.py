"""DEM sensitivity of illumination: the smoothed sun-fraction loss and its adjoint.

``SunLossProblem`` evaluates, over horizon tiles of a site,

    L = mean over valid pixels p and Sun vectors t of F(p, t),

with F the reference's 16-slice finite-disk sun fraction (lightmap.py:32-73)
whose slice clamps are rounded over +-``smoothing_deg`` (``lsh_sun_loss``), and
its gradient with respect to every DEM elevation: the forward march stores the
int32 winning-sample index per (pixel, azimuth); the backward computes
dL/d(horizon) and re-marches each ray with a non-zero gradient to its winning
sample (``lsh_horizon_adjoint``), scattering into the sample's four bilinear
nodes and the pixel's observer-terrain nodes through deterministic 64-bit
fixed-point atomics.

``HorizonSunLoss`` is the ``torch.autograd.Function`` around it (PyTorch is
used only here): ``loss = HorizonSunLoss.apply(problem, primary, regional)``,
then ``loss.backward()`` gives dL/dDEM on the elevation tensors.
"""

from __future__ import annotations

import ctypes
import functools
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import DeviceBuffer, check
from .geometry import DemGrid
from .lighting import SUN_HALF_ANGLE_DEG, validate_vectors

PIXELS = 128 * 128


class SunLossProblem:
    """Forward/backward of the smoothed sun-fraction loss over tiles of one site."""

    def __init__(self, dems: Sequence[DemGrid], tiles: Sequence[tuple[int, int]], sun_vectors_m, *,
                 azimuth_count: int = 720, smoothing_deg: float = 0.02,
                 observer_elevation_m: float = 0.0, session=None, device_id: int = 0,
                 fixed_bits: int = 32):
        from .hip_backend import HipSession

        if not 1 <= len(dems) <= 4:
            raise ValueError("one to four DEMs")
        self.session = session if session is not None else HipSession(device_id)
        self.dems = list(dems)
        self.tiles = [(int(c), int(r)) for c, r in tiles]
        self.sun = validate_vectors(sun_vectors_m)
        self.N = int(azimuth_count)
        self.smoothing_deg = float(smoothing_deg)
        self.observer = float(observer_elevation_m)
        self.fixed_scale = float(2 ** int(fixed_bits))
        s = self.session
        n = len(self.tiles)
        self._sun = DeviceBuffer.from_host(s, self.sun)
        self._horizons = DeviceBuffer.empty(s, (n, PIXELS, self.N), np.float32)
        self._winners = DeviceBuffer.empty(s, (n, PIXELS, self.N), np.int32)
        self._grad_h = DeviceBuffer.empty(s, (n, PIXELS, self.N), np.float32)
        self._loss = DeviceBuffer.empty(s, (n, PIXELS), np.float64)
        self._segments = [None] * n
        self._tile_arr = np.ascontiguousarray(np.array(self.tiles, dtype=np.int32).reshape(-1, 2))
        primary = self.dems[0]
        valid = [min(128, primary.width - c) * min(128, primary.height - r) for c, r in self.tiles]
        self.valid_pixels = int(sum(valid))
        self.seconds: dict[str, float] = {}

    # -- inputs ---------------------------------------------------------------------------
    def set_elevations(self, elevations: Sequence[np.ndarray]) -> None:
        """New elevation values for every DEM (same shapes and georeferencing)."""
        if len(elevations) != len(self.dems):
            raise ValueError("one elevation array per DEM")
        new = []
        for dem, z in zip(self.dems, elevations):
            z = np.ascontiguousarray(z, dtype=np.float32)
            if z.shape != dem.elevation_m.shape:
                raise ValueError("elevation shape does not match the DEM")
            new.append(DemGrid(z, dem.geo_transform, dem.projection))
        for dem in self.dems:
            self.session.forget(dem)
        self.dems = new

    # -- passes ---------------------------------------------------------------------------
    def _ids(self) -> np.ndarray:
        return np.array([self.session.dem_id(d) for d in self.dems], dtype=np.int32)

    def _slot(self, buf: DeviceBuffer, i: int, dtype, shape):
        view = DeviceBuffer.__new__(DeviceBuffer)
        view.session, view.dtype, view.shape = buf.session, np.dtype(dtype), shape
        view.nbytes = int(np.prod(shape)) * view.dtype.itemsize
        view.ptr = buf.ptr + i * view.nbytes
        view.free = lambda: None
        return view

    def forward(self) -> float:
        """Horizons (degrees) with winners for every tile, then the loss."""
        from ._lib import Event

        s = self.session
        ids = self._ids()
        e0 = Event(s).record()
        for i, (c, r) in enumerate(self.tiles):
            self._segments[i] = s.segments_device(self.dems, tile_column=c, tile_row=r,
                                                  azimuth_count=self.N,
                                                  observer_elevation_m=self.observer,
                                                  out=self._segments[i])
            check(s.lib.lsh_horizon_forward(
                s.handle, ids.ctypes.data, len(ids), self._segments[i].ptr, self.N, c, r,
                self.observer, _lib.HORIZON_DEGREES,
                self._slot(self._horizons, i, np.float32, (PIXELS, self.N)).ptr,
                self._slot(self._winners, i, np.int32, (PIXELS, self.N)).ptr),
                "lsh_horizon_forward")
        e1 = Event(s).record()
        check(s.lib.lsh_sun_loss(s.handle, ids[0], self._tile_arr.ctypes.data, len(self.tiles),
                                 self._horizons.ptr, self.N, self._sun.ptr, len(self.sun),
                                 SUN_HALF_ANGLE_DEG, self.smoothing_deg, self._loss.ptr, None),
              "lsh_sun_loss")
        e2 = Event(s).record()
        s.synchronize()
        self.seconds["forward_march"] = e0.elapsed_ms(e1) / 1000.0
        self.seconds["forward_loss"] = e1.elapsed_ms(e2) / 1000.0
        return float(self._loss.to_host().sum() / (self.valid_pixels * len(self.sun)))

    def backward(self, *, sample_cells: bool = False):
        """dL/d(elevation) per DEM (float64 arrays shaped like the DEMs).

        With ``sample_cells`` also returns int32 [tile][16384][N] winning cells
        (pass << 28 | row << 14 | column, -1 where the ray was skipped).
        """
        from ._lib import Event

        s = self.session
        ids = self._ids()
        grads = [DeviceBuffer.empty(s, d.elevation_m.shape, np.int64) for d in self.dems]
        for g in grads:
            g.zero()
        cells = DeviceBuffer.empty(s, (len(self.tiles), PIXELS, self.N), np.int32) if sample_cells else None
        pointers = (ctypes.c_void_p * len(grads))(*[g.ptr for g in grads])
        e0 = Event(s).record()
        check(s.lib.lsh_sun_loss(s.handle, ids[0], self._tile_arr.ctypes.data, len(self.tiles),
                                 self._horizons.ptr, self.N, self._sun.ptr, len(self.sun),
                                 SUN_HALF_ANGLE_DEG, self.smoothing_deg, None, self._grad_h.ptr),
              "lsh_sun_loss")
        e1 = Event(s).record()
        for i, (c, r) in enumerate(self.tiles):
            check(s.lib.lsh_horizon_adjoint(
                s.handle, ids.ctypes.data, len(ids), self._segments[i].ptr, self.N, c, r,
                self.observer, self._slot(self._winners, i, np.int32, (PIXELS, self.N)).ptr,
                self._slot(self._grad_h, i, np.float32, (PIXELS, self.N)).ptr,
                ctypes.cast(pointers, ctypes.c_void_p), self.fixed_scale,
                self._slot(cells, i, np.int32, (PIXELS, self.N)).ptr if cells is not None else None),
                "lsh_horizon_adjoint")
        e2 = Event(s).record()
        s.synchronize()
        self.seconds["backward_loss"] = e0.elapsed_ms(e1) / 1000.0
        self.seconds["backward_march"] = e1.elapsed_ms(e2) / 1000.0
        norm = 1.0 / (self.fixed_scale * self.valid_pixels * len(self.sun))
        out = [g.to_host().astype(np.float64) * norm for g in grads]
        for g in grads:
            g.free()
        if cells is not None:
            c = cells.to_host()
            cells.free()
            return out, c
        return out

    def horizons(self) -> np.ndarray:
        """Degrees [tile][16384][N] of the last forward."""
        return self._horizons.to_host()

    def grad_horizons(self) -> np.ndarray:
        """d(sum of F) / d(horizon degrees) [tile][16384][N] of the last backward."""
        return self._grad_h.to_host()

    def close(self) -> None:
        for b in [self._sun, self._horizons, self._winners, self._grad_h, self._loss, *self._segments]:
            if b is not None:
                b.free()


@functools.lru_cache(maxsize=None)
def _autograd_function():
    import torch

    class HorizonSunLoss(torch.autograd.Function):
        """loss = HorizonSunLoss.apply(problem, *elevations) (float64 scalar tensor)."""

        @staticmethod
        def forward(ctx, problem: SunLossProblem, *elevations):
            problem.set_elevations([e.detach().to("cpu", torch.float32).numpy() for e in elevations])
            loss = problem.forward()
            ctx.problem = problem
            ctx.meta = [(e.dtype, e.device) for e in elevations]
            return torch.tensor(loss, dtype=torch.float64, device=elevations[0].device)

        @staticmethod
        def backward(ctx, grad_output):
            grads = ctx.problem.backward()
            out = [torch.from_numpy(g).to(device=device, dtype=dtype) * grad_output.to(device)
                   for g, (dtype, device) in zip(grads, ctx.meta)]
            return (None, *out)

    return HorizonSunLoss


def __getattr__(name):
    if name == "HorizonSunLoss":
        return _autograd_function()
    raise AttributeError(name)
