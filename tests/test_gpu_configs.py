"""GPU parity at the workloads of the five benchmark configurations.

Every case compares the HIP path with the float64 oracle on the same seeded
synthetic inputs, at the configuration's own DEM sizes (nothing shrunk):

1. one 1024^2 @20 m DEM (seed 0), tile (512, 512): all 16384 px x 1440 az;
2. the full site, 2048^2 @5 m (seed 1) + 4096^2 @60 m (seed 2): a corner, two
   edge tiles and a centre tile, 1024 seeded pixels x 1440 az each;
3. three of the 48 landing sites (1, 4 and 12 tiles) sharing the regional DEM,
   marched by the persistent batched kernel: every tile bit-identical to the
   per-tile path, and one tile per site against the oracle (256 px);
4. 4380 Sun + 4380 Earth vectors (2-hour steps over 365 days) and their PSR
   reduction streamed over two full-site tiles: sun fraction, Earth margin,
   Earth visibility and the PSR mask against the oracle's lighting of the same
   horizons.

Bars: max |delta horizon| <= 0.01 deg with >= 99.9 % of bins within 1e-3 deg;
sun fraction within 1e-3 of the oracle's float64 frame and az/el feeding the
reference's float32 slice walk (the float64 walk's difference is printed: the
float32 bin position alone moves a fraction by up to ~1e-3 where the horizon is
steep); Earth margin within 1e-4 deg; visibility within one
time step; PSR identical except on pixels whose best margin lies within 1e-3
deg of the limb threshold. Each case prints bins, max |delta| and the fraction
within 1e-3 deg. The adjoint configuration (5) is in test_gpu_adjoint.py.
"""

from __future__ import annotations

import functools
import time

import numpy as np
import pytest

import oracle
from lunarscout_amd import synthetic

pytestmark = pytest.mark.gpu

MAX_DEG = 0.01
BULK_DEG = 1e-3
BULK_FRACTION = 0.999


@functools.lru_cache(maxsize=None)
def session():
    from lunarscout_amd.hip_backend import HipSession

    return HipSession(0)


@functools.lru_cache(maxsize=None)
def full_site():
    return synthetic.site_dems(synthetic.FULL_SITE)


def _parity(gpu_slopes, oracle_slopes, label):
    g = np.degrees(np.arctan(np.asarray(gpu_slopes, np.float64)))
    o = np.degrees(np.arctan(np.asarray(oracle_slopes, np.float64)))
    finite = np.isfinite(o)
    np.testing.assert_array_equal(np.isfinite(g), finite, err_msg=f"{label}: -inf pattern")
    err = np.abs(g - o)[finite]
    worst = float(err.max()) if err.size else 0.0
    within = float(np.mean(err <= BULK_DEG)) if err.size else 1.0
    print(f"{label}: bins {err.size}, max |delta| {worst:.3g} deg, within 1e-3 deg {within:.6f}",
          flush=True)
    assert worst <= MAX_DEG, label
    assert within >= BULK_FRACTION, label
    return worst, within


def _oracle_tile(dems, tile, pixels):
    t0 = time.time()
    segs = oracle.segments(dems, tile_column=tile[0], tile_row=tile[1], tile_width=128,
                           azimuth_count=1440)
    pyramids = [oracle.OraclePyramid(d) for d in dems]
    ref, _ = oracle.horizon_tile(pyramids, segs, tile_column=tile[0], tile_row=tile[1],
                                 pixels=pixels, float64=True)
    print(f"  oracle: {len(pixels)} px x 1440 az x {len(dems)} passes in {time.time() - t0:.1f} s "
          f"on {oracle.cpu_threads()} threads", flush=True)
    return ref


def test_config1_whole_tile_against_the_oracle():
    dems = [synthetic.fbm_crater_dem(1024, 20.0, 0)]
    tile = (512, 512)
    got = session().horizon_tile(dems, tile_column=tile[0], tile_row=tile[1])
    pixels = np.arange(16384, dtype=np.int32)
    ref = _oracle_tile(dems, tile, pixels)
    _parity(got, ref, "config 1, tile (512, 512), all 16384 px")


@pytest.mark.parametrize("tile", [(0, 0), (1920, 896), (640, 1920), (1024, 1024)],
                         ids=["corner", "right-edge", "bottom-edge", "centre"])
def test_config2_full_site_tiles_against_the_oracle(tile):
    dems = full_site()
    got = session().horizon_tile(dems, tile_column=tile[0], tile_row=tile[1])
    rng = np.random.default_rng(tile[0] * 7 + tile[1])
    pixels = np.sort(rng.choice(16384, 1024, replace=False)).astype(np.int32)
    ref = _oracle_tile(dems, tile, pixels)
    _parity(got[pixels], ref, f"config 2, tile {tile}, 1024 px")


def test_config3_batched_landing_sites_against_plain_and_oracle():
    from lunarscout_amd._lib import DeviceBuffer
    from lunarscout_amd.batch import HorizonBatch, Site

    s = session()
    regional = full_site()[1]
    sites = {site.index: site for site in synthetic.landing_sites()}
    chosen = [sites[13], sites[3], sites[45]]  # 1, 4 and 12 tiles
    patch_dems = [synthetic.landing_site_dem(site) for site in chosen]
    work = [(i, (k % site.tiles_x) * 128, (k // site.tiles_x) * 128)
            for i, site in enumerate(chosen) for k in range(site.tile_count)]
    batch = HorizonBatch(s, [Site((d, regional)) for d in patch_dems], max_tiles=len(work))
    out = DeviceBuffer.empty(s, (len(work), 16384, 1440), np.float32)
    batch.run(work, out, use_graph=True)
    batched = out.to_host()
    batch.close()
    out.free()
    identical = 0
    for slot, (i, c, r) in enumerate(work):
        plain = s.horizon_tile([patch_dems[i], regional], tile_column=c, tile_row=r)
        same = (plain == batched[slot]) | (np.isneginf(plain) & np.isneginf(batched[slot]))
        assert same.all(), (chosen[i].index, c, r)
        identical += 1
    print(f"config 3: {identical} tiles of sites {[site.index for site in chosen]} "
          "batched == plain bit for bit", flush=True)
    rng = np.random.default_rng(3)
    for i, site in enumerate(chosen):
        slot = next(k for k, w in enumerate(work) if w[0] == i)
        _, c, r = work[slot]
        pixels = np.sort(rng.choice(16384, 256, replace=False)).astype(np.int32)
        ref = _oracle_tile([patch_dems[i], regional], (c, r), pixels)
        _parity(batched[slot][pixels], ref,
                f"config 3, site {site.index} ({site.tile_count} tiles, "
                f"{site.resolution_m:.1f} m), tile ({c}, {r}), 256 px")


def test_config4_year_of_lighting_against_the_oracle():
    from lunarscout_amd._lib import DeviceBuffer
    from lunarscout_amd.lighting import HipLightingSession, reduce_sun_vectors_for_psr

    s = session()
    dems = full_site()
    primary = dems[0]
    tiles = [(1024, 1024), (0, 1920)]
    horizons = np.stack([s.horizon_tile(dems, tile_column=c, tile_row=r, degrees=True)
                         for c, r in tiles])
    sun = synthetic.sun_vectors(primary, 4380, 2.0, seed=5)
    earth = synthetic.earth_vectors(4380, 2.0, seed=6)
    reduced, _ = reduce_sun_vectors_for_psr(primary, sun)
    lit = HipLightingSession(s)
    hz = DeviceBuffer.from_host(s, horizons)
    out = lit.lighting_tiles(primary, tiles, hz, sun=sun, earth=earth, psr_vectors=reduced,
                             sun_fraction=True, sun_mean=True, earth_margin=True,
                             earth_visible=True, psr=True, psr_margin=True)
    fraction = out.sun_fraction.to_host()
    margin = out.earth_margin.to_host()
    visible = out.earth_visible.to_host()
    psr = out.psr.to_host()
    out.free()
    hz.free()
    threshold = -float(np.float32(0.545)) / 2
    for i, (c, r) in enumerate(tiles):
        rows = horizons[i].reshape(128, 128, -1)
        # the reference's slice walk is float32 (lightmap.py:32-73): compare with the
        # oracle's float64 frame and az/el feeding that walk, and report the float64 walk
        ref = oracle.lighting_patch(primary, rows, sun, tile_y=r, tile_x=c, float32_slices=True)
        err = np.abs(fraction[i].reshape(-1, 128, 128) - ref["fraction"])
        ref64 = oracle.lighting_patch(primary, rows, sun, tile_y=r, tile_x=c)["fraction"]
        err64 = float(np.abs(fraction[i].reshape(-1, 128, 128) - ref64).max())
        del ref64
        eref = oracle.lighting_patch(primary, rows, earth, tile_y=r, tile_x=c)
        merr = np.abs(margin[i].reshape(-1, 128, 128) - eref["margin"])
        verr = np.abs(visible[i].reshape(128, 128) - eref["visible_fraction"])
        mask, best = oracle.psr_patch(primary, rows, reduced, tile_y=r, tile_x=c)
        differ = psr[i].reshape(128, 128) != mask
        near = np.abs(best - threshold) <= 1e-3
        print(f"config 4, tile {(c, r)}: {fraction[i].size} sun px-times, max |d fraction| "
              f"{err.max():.3g} (float64 slice walk: {err64:.3g}); {margin[i].size} Earth "
              f"px-times, max |d margin| {merr.max():.3g} "
              f"deg, max |d visibility| {verr.max():.3g}; {len(reduced)} PSR vectors, PSR px "
              f"{int((mask == 255).sum())}, differing {int(differ.sum())} "
              f"(near threshold {int((differ & near).sum())})", flush=True)
        assert err.max() <= 1e-3
        assert merr.max() <= 1e-4
        assert verr.max() <= 1.0 / len(earth) + 1e-6
        assert not np.any(differ & ~near)
