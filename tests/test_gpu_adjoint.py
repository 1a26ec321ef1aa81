"""The adjoint: dL/dDEM of the smoothed sun-fraction loss.

* the loss and dL/d(horizon) against a pure-PyTorch autograd reference of the
  same smoothed 16-slice model (float64), on a small two-DEM case;
* dL/dDEM of the adjoint configuration (512^2 @10 m + 1024^2 @60 m, 16 tiles x
  720 azimuths, 2190 Sun vectors) against central finite differences with
  h = 5 cm of the float64 oracle under the adjoint's convention -- sample
  positions and culls held at the forward's (the march's adaptive steps also
  move samples continuously with the heights; the winning-sample adjoint does
  not differentiate that): for each probed node the affected pixels (those
  whose winning sample touches the node, and the pixel the node carries) are
  re-evaluated by the oracle with the node moved by +-h, and their loss
  recomputed in float64; at least 200 tie-free nodes within 1e-3 relative;
* run-to-run bit-identical gradients (fixed-point scatter), and the
  torch.autograd.Function wrapper.
"""

from __future__ import annotations

import functools
import time

import numpy as np
import pytest

import oracle
from lunarscout_amd import synthetic
from lunarscout_amd.geometry import DemGrid

pytestmark = pytest.mark.gpu

N = 720


@functools.lru_cache(maxsize=None)
def _case():
    from lunarscout_amd.adjoint import SunLossProblem

    dems = [synthetic.fbm_crater_dem(256, 10.0, 7), synthetic.fbm_crater_dem(512, 60.0, 8)]
    tiles = [(0, 0), (128, 0), (0, 128), (128, 128)]
    sun = synthetic.sun_vectors(dems[0], 365, 4.0, seed=9)
    problem = SunLossProblem(dems, tiles, sun, azimuth_count=N)
    loss = problem.forward()
    grads, cells = problem.backward(sample_cells=True)
    return problem, loss, grads, cells


def test_loss_and_horizon_gradient_match_a_pytorch_reference():
    import torch

    problem, loss, _, _ = _case()
    horizons = problem.horizons()
    grad_h = problem.grad_horizons()
    rng = np.random.default_rng(0)
    half, max_photons = oracle._sun_disk(0.27)
    s = problem.smoothing_deg
    for tile_index, (tx, ty) in enumerate(problem.tiles[:2]):
        pixels = rng.choice(16384, 24, replace=False)
        for p in pixels:
            row, column = ty + p // 128, tx + p % 128
            az, el = oracle.azimuth_elevation(problem.dems[0], row, column, problem.sun)
            h = torch.tensor(np.maximum(horizons[tile_index, p].astype(np.float64), -90.0),
                             requires_grad=True)
            # the slice walk's bins and float32 fractions (index work, no gradient)
            # N bins: x N/360 bins per degree, half a bin, slices 0.135 * (N / 1440) bins apart
            per_degree = np.float32(N) / np.float32(360.0)
            half_bucket = np.float32(0.5) / per_degree
            slice_step = np.float32(0.135) * (per_degree * np.float32(0.25))
            left_pos = ((az.astype(np.float32) - np.float32(0.27)) - half_bucket) * per_degree
            left = np.trunc(left_pos).astype(np.int64)
            frac = (left_pos - left.astype(np.float32)).astype(np.float32)
            left = np.where(left < 0, left + N, np.where(left >= N, left - N, left))
            right = np.where(left + 1 >= N, 0, left + 1)
            lo = left.copy()
            el_t = torch.tensor(el.astype(np.float32).astype(np.float64))
            total = torch.zeros((), dtype=torch.float64)
            for i in range(16):
                c = float(half[i])
                f = torch.tensor(frac.astype(np.float64))
                horizon = f * (h[torch.tensor(right)] - h[torch.tensor(lo)]) + h[torch.tensor(lo)]
                x = el_t + c - horizon
                W = 2.0 * c
                value = torch.where(x <= -s, torch.zeros_like(x),
                        torch.where(x < s, (x + s) ** 2 / (4 * s),
                        torch.where(x <= W - s, x,
                        torch.where(x < W + s, W - (W + s - x) ** 2 / (4 * s), torch.full_like(x, W)))))
                total = total + value.sum()
                frac = (frac + slice_step).astype(np.float32)
                step = frac >= np.float32(1.0)
                lo = np.where(step, right, lo)
                right = np.where(step, np.where(right + 1 >= N, 0, right + 1), right)
                frac = np.where(step, (frac - np.float32(1.0)).astype(np.float32), frac)
            total = total / float(max_photons)
            total.backward()
            got = grad_h[tile_index, p].astype(np.float64)
            want = h.grad.numpy()
            scale = max(np.abs(want).max(), 1e-12)
            assert np.abs(got - want).max() <= 1e-5 * scale + 1e-7, (tile_index, p)
    assert 0.0 < loss < 1.0


def _perturbed(dem: DemGrid, x: int, y: int, delta: float) -> tuple[DemGrid, float]:
    """The DEM with node (x, y) moved by delta, and the move float32 storage made."""
    z = dem.elevation_m.copy()
    z[y, x] = np.float32(np.float64(z[y, x]) + delta)
    return DemGrid(z, dem.geo_transform, dem.projection), float(np.float64(z[y, x]) -
                                                               np.float64(dem.elevation_m[y, x]))


def _affected_pixels(problem, cells, nodes):
    """{(dem, x, y): sorted (tile, pixel) pairs whose loss can move with that node}.

    A pixel is affected when the winning sample of one of its rays interpolates
    the node (the sample's cell is one of the four cells around it) or, for the
    primary DEM, when the node is the pixel's own observer terrain.
    """
    code_to_nodes: dict[int, list] = {}
    for node in nodes:
        d, x, y = node
        for cy in (y - 1, y):
            for cx in (x - 1, x):
                code_to_nodes.setdefault((d << 28) | (cy << 14) | cx, []).append(node)
    codes = np.fromiter(code_to_nodes.keys(), dtype=np.int64)
    out = {node: set() for node in nodes}
    for t, (tx, ty) in enumerate(problem.tiles):
        c = cells[t].astype(np.int64)
        hit = np.isin(c, codes)
        pix, az = np.nonzero(hit)
        for p, code in set(zip(pix.tolist(), c[pix, az].tolist())):
            for node in code_to_nodes[code]:
                out[node].add((t, p))
        for node in nodes:
            d, x, y = node
            if d == 0 and tx <= x < tx + 128 and ty <= y < ty + 128:
                out[node].add((t, (y - ty) * 128 + (x - tx)))
    return {node: sorted(v) for node, v in out.items()}


def _fd_loss(problem, dems, affected, segments, pyramids):
    """Float64 loss of the affected pixels: the oracle marches the base DEMs and
    evaluates every sample on ``dems`` (the adjoint's frozen-schedule convention)."""
    total = 0.0
    by_tile: dict[int, list[int]] = {}
    for t, p in affected:
        by_tile.setdefault(t, []).append(p)
    for t, pixels in by_tile.items():
        tx, ty = problem.tiles[t]
        slopes = oracle.horizon_tile_frozen(pyramids, dems, segments[t], tile_column=tx,
                                            tile_row=ty, pixels=np.array(pixels, np.int32))
        degrees = np.degrees(np.arctan(slopes))
        degrees[~np.isfinite(slopes)] = -90.0
        total += oracle.smoothed_sun_loss(dems[0], degrees, problem.sun, tile_y=ty, tile_x=tx,
                                          pixels=pixels, smoothing_deg=problem.smoothing_deg).sum()
    return total


# The adjoint configuration: 512^2 @10 m (seed 7) + 1024^2 @60 m (seed 8), 16 tiles
# x 720 azimuths, 2190 Sun vectors at 4-hour steps over 365 days (seed 10).
FD_STEP_M = 0.05
FD_SMOOTHING_DEG = 0.08
FD_NODES = 200
FD_MAX_AFFECTED = 24


@functools.lru_cache(maxsize=None)
def _config5():
    from lunarscout_amd.adjoint import SunLossProblem

    dems = synthetic.site_dems(synthetic.ADJOINT_SITE)
    tiles = [(x, y) for y in range(0, 512, 128) for x in range(0, 512, 128)]
    sun = synthetic.sun_vectors(dems[0], 2190, 4.0, seed=10)
    problem = SunLossProblem(dems, tiles, sun, azimuth_count=N, smoothing_deg=FD_SMOOTHING_DEG)
    loss = problem.forward()
    grads, cells = problem.backward(sample_cells=True)
    return problem, loss, grads, cells


def test_adjoint_matches_finite_differences_at_5_cm_on_200_nodes():
    """dL/dDEM of the config-5 problem against central differences with h = 5 cm.

    The loss's slice clamps are rounded over +-0.08 deg (at most the narrowest
    slice half-height, 0.094 deg): a 5 cm move of a node changes the horizon of
    a ray that samples it r metres away by about 2.9 / r degrees, so rays beyond
    a few hundred metres stay inside the rounding and 5 cm is a linear step
    there. A node is used when it is tie-free and linear at that step: its
    forward and backward differences at h = 5 cm agree to 2e-4 of the central
    one (curvature over the step, a winner switch or a slice edge crossed inside
    it, or a tie at the base state -- a kink that central differences alone
    average away -- breaks that). Nodes
    are drawn at random among interior nodes of both DEMs whose gradient is
    at least 1e5 fixed-point steps of the backward's scatter, with at most 24
    affected pixels each (so the oracle finishes in seconds); at least 200 must
    pass and agree with the adjoint to 1e-3 relative.
    """
    problem, _, grads, cells = _config5()
    dems = problem.dems
    t0 = time.time()
    segments = [oracle.segments(dems, tile_column=tx, tile_row=ty, tile_width=128, azimuth_count=N)
                for tx, ty in problem.tiles]
    pyramids = [oracle.OraclePyramid(d) for d in dems]
    norm = problem.valid_pixels * len(problem.sun)
    rng = np.random.default_rng(5)
    candidates = []
    for dem_index, count in ((0, 2000), (1, 1500)):
        g = grads[dem_index]
        h_, w_ = g.shape
        interior = np.zeros_like(g, dtype=bool)
        interior[1:h_ - 2, 1:w_ - 2] = True
        # the backward sums contributions in 64-bit fixed point with steps of
        # 1 / fixed_scale of the un-normalised sum: a node is compared only where
        # its gradient is at least 1e5 such steps (quantisation below 1e-4
        # relative even with tens of contributions)
        units = np.abs(g) * norm * problem.fixed_scale
        nz = np.argwhere((units >= 1e5) & interior)
        pick = nz[rng.choice(len(nz), min(count, len(nz)), replace=False)]
        candidates += [(dem_index, int(x), int(y)) for y, x in pick]
    order = rng.permutation(len(candidates))
    candidates = [candidates[i] for i in order]
    affected = _affected_pixels(problem, cells, candidates)
    print(f"config 5: {len(candidates)} candidate nodes, affected sets in {time.time() - t0:.1f} s",
          flush=True)
    accepted = {0: 0, 1: 0}
    rejected_nonlinear = skipped = 0
    worst = 0.0
    for node in candidates:
        if sum(accepted.values()) >= FD_NODES:
            break
        d, x, y = node
        pixels = affected[node]
        if not pixels or len(pixels) > FD_MAX_AFFECTED:
            skipped += 1
            continue
        values, moves = [], []
        for sign in (1.0, 0.0, -1.0):
            moved = list(dems)
            moved[d], actual = _perturbed(dems[d], x, y, sign * FD_STEP_M)
            values.append(_fd_loss(problem, moved, pixels, segments, pyramids))
            moves.append(actual)
        central = (values[0] - values[2]) / (moves[0] - moves[2]) / norm
        forward = (values[0] - values[1]) / (moves[0] - moves[1]) / norm
        backward = (values[1] - values[2]) / (moves[1] - moves[2]) / norm
        # one-sided differences that disagree mean curvature over the step or a
        # kink at the base (a winner tie): not tie-free and linear at 5 cm
        if abs(forward - backward) > 2e-4 * max(abs(central), 1e-30):
            rejected_nonlinear += 1
            continue
        fd = [central]
        rel = abs(grads[d][y, x] - fd[0]) / max(abs(fd[0]), 1e-30)
        worst = max(worst, rel)
        accepted[d] += 1
        if sum(accepted.values()) % 25 == 0:
            print(f"  {sum(accepted.values())} nodes checked, worst rel {worst:.2e} "
                  f"({time.time() - t0:.0f} s)", flush=True)
        assert rel <= 1e-3, (node, grads[d][y, x], fd[0])
    print(f"config 5: {sum(accepted.values())} nodes within 1e-3 of central differences at "
          f"h = 5 cm (primary {accepted[0]}, regional {accepted[1]}), worst rel {worst:.2e}; "
          f"rejected as non-linear or tied at 5 cm: {rejected_nonlinear}; skipped (more than "
          f"{FD_MAX_AFFECTED} affected px): {skipped}; {time.time() - t0:.0f} s", flush=True)
    assert sum(accepted.values()) >= FD_NODES
    assert accepted[0] > 0 and accepted[1] > 0


def test_gradients_are_bit_identical_run_to_run():
    problem, _, grads, _ = _case()
    again = problem.backward()
    for a, b in zip(grads, again):
        assert a.tobytes() == b.tobytes()


def test_autograd_function_matches_the_problem():
    import torch
    from lunarscout_amd.adjoint import HorizonSunLoss, SunLossProblem

    dems = [synthetic.fbm_crater_dem(256, 10.0, 7), synthetic.fbm_crater_dem(512, 60.0, 8)]
    sun = synthetic.sun_vectors(dems[0], 64, 4.0, seed=9)
    problem = SunLossProblem(dems, [(128, 128)], sun, azimuth_count=N)
    z = [torch.tensor(d.elevation_m, requires_grad=True) for d in dems]
    loss = HorizonSunLoss.apply(problem, *z)
    loss.backward()
    reference = SunLossProblem(dems, [(128, 128)], sun, azimuth_count=N)
    assert abs(reference.forward() - loss.item()) <= 1e-12
    want = reference.backward()
    for tensor, w in zip(z, want):
        np.testing.assert_allclose(tensor.grad.numpy(), w.astype(np.float32), rtol=1e-6, atol=1e-30)
    problem.close()
    reference.close()
