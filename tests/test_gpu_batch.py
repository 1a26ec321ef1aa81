"""The batched persistent march equals the plain per-tile march bit for bit.

Every kernel shape, with and without the hipGraph, over tiles of several sites
(different primary DEMs and resolutions sharing one regional DEM), including
partial edge tiles and a replay with a different tile list.
"""

from __future__ import annotations

import functools

import numpy as np
import pytest

from lunarscout_amd import synthetic
from lunarscout_amd._lib import DeviceBuffer
from lunarscout_amd.geometry import DemGrid

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=None)
def _case():
    from lunarscout_amd.batch import Site
    from lunarscout_amd.hip_backend import HipSession

    session = HipSession(0)
    regional = synthetic.fbm_crater_dem(1024, 60.0, 2)
    sites = []
    for i, (h, w, res) in enumerate(((256, 384, 10.0), (200, 130, 20.0), (128, 128, 5.0))):
        dem = synthetic.fbm_crater_dem(max(h, w), res, 100 + i,
                                       centre_xy=(synthetic.site_xy()[0] + 4000.0 * i,
                                                  synthetic.site_xy()[1] - 3000.0 * i))
        dem = DemGrid(np.ascontiguousarray(dem.elevation_m[:h, :w]), dem.geo_transform, dem.projection)
        sites.append(Site((dem, regional), observer_elevation_m=0.5 * i))
    return session, sites


def _plain(session, site, column, row):
    return session.horizon_tile(list(site.dems), tile_column=column, tile_row=row,
                                observer_elevation_m=site.observer_elevation_m)


def _check(session, sites, batch, tiles, out, **kw):
    batch.run(tiles, out, **kw)
    session.synchronize()
    got = out.to_host()
    for i, (s, c, r) in enumerate(tiles):
        want = _plain(session, sites[s], c, r)
        same = (got[i] == want) | (np.isneginf(got[i]) & np.isneginf(want))
        print(f"tile {i} site {s} ({c},{r}) {kw}: identical {same.mean():.6f}", flush=True)
        assert same.all()


def test_batch_equals_plain_for_every_shape_and_with_the_graph():
    from lunarscout_amd.batch import HorizonBatch, shapes

    session, sites = _case()
    batch = HorizonBatch(session, sites, max_tiles=4)
    out = DeviceBuffer.empty(session, (4, 16384, 1440), np.float32)
    tiles = [(0, 256, 128), (1, 0, 128), (2, 0, 0)]
    for config in range(len(shapes())):
        _check(session, sites, batch, tiles[:2], out, config=config, use_graph=False)
    _check(session, sites, batch, tiles, out, config=0, use_graph=True)
    # replay of the captured graph with a different tile list of the same length
    _check(session, sites, batch, [(0, 0, 0), (1, 128, 0), (0, 128, 128)], out, config=0,
           use_graph=True)
    # a persistent grid smaller than the device, so queues are stolen from
    _check(session, sites, batch, tiles, out, config=2, workgroups=64, use_graph=True)
    batch.close()
    out.free()


def test_autotune_returns_a_measured_shape(tmp_path, monkeypatch):
    from lunarscout_amd.batch import HorizonBatch, autotune

    monkeypatch.setenv("LSH_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    session, sites = _case()
    batch = HorizonBatch(session, sites, max_tiles=2, azimuth_count=360)
    out = DeviceBuffer.empty(session, (2, 16384, 360), np.float32)
    result = autotune(batch, [(0, 128, 128), (2, 0, 0)], out, key="test", configs=[0, 2],
                      workgroup_factors=(1.0,))
    assert result["config"] in (0, 2) and len(result["trials"]) == 2 and not result["cached"]
    again = autotune(batch, [(0, 128, 128)], out, key="test")
    assert again["cached"] and again["config"] == result["config"]
    batch.close()
    out.free()
