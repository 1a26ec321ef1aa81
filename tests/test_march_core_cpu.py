"""The kernel's march core (CPU build) against the oracle on constructed rays.

The GPU kernel is checked bit for bit against the CPU build of the same header
(tests/test_gpu_horizon.py), so a deviation from the reference that both share
is invisible there. These rays are built to land exactly on a cell's minimum
edge while the tangent points out of the cell (the exit distance is 0), which
is where the reference takes the fallback distance (cuda_backend.py:752-765,
794-801 for cull and nodata skips, :636-640 for the diagnostic cap): the skip
is fallback + 1e-4 km and the capped step fallback + 1e-6 km. The oracle
(oracle/lsh_oracle.cpp hierarchy_maximum) follows those lines; the kernel core
must produce the same trace row for row.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from lunarscout_amd.geometry import DemGrid, ProjectionParameters
from tests import kernel_cpu

RES = 20.0


def _pyramid(elevation: np.ndarray) -> oracle.OraclePyramid:
    transform = np.array([0.0, RES, 0.0, 0.0, 0.0, -RES], dtype=np.float64)
    dem = DemGrid(np.ascontiguousarray(elevation, dtype=np.float32), transform,
                  ProjectionParameters.south_polar())
    return oracle.OraclePyramid(dem)


def _segment(start_x: float, a1: float, *, a2=0.0, s_start=0.6, s_end=1.5) -> np.ndarray:
    """A ray along -x/+x from (start_x, 100.5); straight rays have true distance = s (km) * 1000."""
    seg = np.zeros(18, dtype=np.float32)
    seg[0], seg[1] = start_x, 100.5
    seg[2], seg[3] = start_x, 100.5
    seg[4], seg[5] = a1, a2          # pixels per km (and per km^2) along x
    seg[12], seg[13] = s_start, s_end
    seg[14] = s_start                # chord start (km)
    seg[15] = 1.0                    # c1: chord metres = planar metres
    return seg


def _both(segment, pyramid, *, observer_z, initial=-np.inf, cap=False):
    mode = oracle.CAP if cap else 0
    ref, ref_rows = oracle.march_ray(segment, pyramid, observer_z=observer_z,
                                     radius=1737400.0, map_resolution=RES, pass_index=0,
                                     initial=initial, mode=mode, trace_capacity=20000)
    got, got_rows = kernel_cpu.march_segment(segment, pyramid, observer_z=observer_z,
                                             initial=initial, mode=6, cap=cap)
    ref_rows = ref_rows[ref_rows[:, 11] != 3]  # the oracle also records the ray leaving a level
    return ref, ref_rows, got, got_rows


def _assert_same_trace(ref_rows, got_rows):
    assert len(ref_rows) == len(got_rows), (len(ref_rows), len(got_rows))
    for column in (0, 1, 5, 6, 10):  # s, true distance, pixel position, advance
        np.testing.assert_allclose(got_rows[:, column], ref_rows[:, column], rtol=2e-6,
                                   atol=1e-9, err_msg=f"trace column {column}")
    for column in (2, 3, 4, 11):  # level, cell, action
        np.testing.assert_array_equal(got_rows[:, column], ref_rows[:, column])


@pytest.mark.parametrize("start_x", [64.0, 128.0])
def test_cull_skip_from_a_min_edge_moves_fallback_plus_1e4_km(start_x):
    """Every block culls (initial slope 1), the first position sits on a min edge."""
    pyramid = _pyramid(np.zeros((256, 256), dtype=np.float32))
    seg = _segment(start_x, -50.0)
    ref, ref_rows, got, got_rows = _both(seg, pyramid, observer_z=10.0, initial=1.0)
    assert ref_rows[0, 11] == 1 and ref_rows[0, 5] == np.float32(start_x)
    fallback = RES * 0.5 / 1000.0 * (4.0 ** ref_rows[0, 2])
    assert ref_rows[0, 10] == pytest.approx(fallback + 1e-4, rel=1e-6)
    _assert_same_trace(ref_rows, got_rows)
    assert got == pytest.approx(ref)


def test_nodata_skip_from_a_min_edge_moves_fallback_plus_1e4_km():
    elevation = np.full((256, 256), -32768.0, dtype=np.float32)
    pyramid = _pyramid(elevation)
    seg = _segment(64.0, -50.0)
    ref, ref_rows, got, got_rows = _both(seg, pyramid, observer_z=10.0)
    assert ref_rows[0, 11] == 2
    assert ref_rows[0, 10] == pytest.approx(RES * 0.5e-3 + 1e-4, rel=1e-6)
    _assert_same_trace(ref_rows, got_rows)
    assert np.isneginf(ref) and np.isneginf(got)


def test_capped_step_from_a_min_edge_is_fallback_plus_1e6_km():
    """Diagnostic cap: a level-0 sample on a min edge caps its step at fallback + 1e-6 km."""
    rng = np.random.default_rng(3)
    pyramid = _pyramid(rng.normal(0.0, 5.0, (256, 256)).astype(np.float32))
    seg = _segment(64.0, -50.0)
    ref, ref_rows, got, got_rows = _both(seg, pyramid, observer_z=0.0, cap=True)
    first = ref_rows[ref_rows[:, 11] == 4][0]
    assert first[5] == np.float32(64.0)
    assert first[10] == pytest.approx(RES * 0.5e-3 + 1e-6, rel=1e-6)
    _assert_same_trace(ref_rows, got_rows)
    assert got == pytest.approx(ref, rel=1e-6)


def test_edge_rays_on_rough_terrain_match_the_oracle():
    """Rough terrain, rays from integer starts in both directions: slopes and traces agree.

    The rays curve slightly (a2 != 0): a straight ray from an integer start
    would put every level-0 step (one pixel, 1/|tangent| km) exactly on a pixel
    edge again, where float64 rounding of the position decides the cell.
    """
    rng = np.random.default_rng(11)
    elevation = np.cumsum(np.cumsum(rng.normal(0, 1.0, (256, 256)), 0), 1).astype(np.float32)
    pyramid = _pyramid(elevation)
    for start_x, a1 in ((64.0, -50.0), (192.0, -37.5), (16.0, 50.0), (128.0, -12.5)):
        seg = _segment(start_x, a1, a2=0.37, s_start=0.6, s_end=2.4)
        ref, ref_rows, got, got_rows = _both(seg, pyramid, observer_z=float(elevation[100, 64]))
        _assert_same_trace(ref_rows, got_rows)
        assert got == pytest.approx(ref, rel=1e-9, abs=1e-12)


def test_guarded_skips_leave_every_output_and_counter_unchanged():
    """The march with guarded skips (a level-2 bound test that proves the next
    level-0 cells culled, then only their exits) against the same build with
    them off: slopes, winning samples and the traversal counters equal bit for
    bit on a two-DEM site (1024^2 @5 m + 2048^2 @60 m: the regional rays reach
    level 1 beyond 55 km), interior and edge pixels."""
    from lunarscout_amd import synthetic

    dems = synthetic.site_dems(synthetic.FULL_SITE, scale=2)
    pyramids = [oracle.OraclePyramid(d) for d in dems]
    for column, row in [(384, 384), (0, 896)]:
        segs = oracle.segments(dems, tile_column=column, tile_row=row, tile_width=128,
                               azimuth_count=1440)
        pixels = np.arange(0, 16384, 683, dtype=np.int32)
        kw = dict(tile_column=column, tile_row=row, pixels=pixels)
        on, won, con = kernel_cpu.horizon_tile(pyramids, segs, mode=6, **kw)
        off, woff, coff = kernel_cpu.horizon_tile(pyramids, segs, mode=6 | 8, **kw)
        assert con[0] > 0
        np.testing.assert_array_equal(on.view(np.uint32), off.view(np.uint32))
        np.testing.assert_array_equal(won, woff)
        np.testing.assert_array_equal(con, coff)


def test_guarded_skips_with_nodata_holes_leave_every_output_unchanged():
    """As above on the two-DEM site with nodata in both DEMs: scattered NaN pixels,
    a -32768 block and an infinite pixel, so guard windows hold nodata cells and
    whole nodata bounds."""
    from lunarscout_amd import synthetic

    rng = np.random.default_rng(17)
    dems = []
    for d in synthetic.site_dems(synthetic.FULL_SITE, scale=2):
        z = np.array(d.elevation_m, dtype=np.float32)
        h, w = z.shape
        z[rng.random(z.shape) < 0.003] = np.nan
        z[h // 5:h // 5 + 120, w // 2:w // 2 + 140] = -32768.0
        z[h // 3, w // 4] = np.inf
        dems.append(DemGrid(z, d.geo_transform, d.projection))
    pyramids = [oracle.OraclePyramid(d) for d in dems]
    segs = oracle.segments(dems, tile_column=384, tile_row=384, tile_width=128,
                           azimuth_count=1440)
    pixels = np.arange(0, 16384, 683, dtype=np.int32)
    kw = dict(tile_column=384, tile_row=384, pixels=pixels)
    on, won, con = kernel_cpu.horizon_tile(pyramids, segs, mode=6, **kw)
    off, woff, coff = kernel_cpu.horizon_tile(pyramids, segs, mode=6 | 8, **kw)
    assert con[0] > 0
    np.testing.assert_array_equal(on.view(np.uint32), off.view(np.uint32))
    np.testing.assert_array_equal(won, woff)
    np.testing.assert_array_equal(con, coff)
