"""lsh_lighting_tiles_static (no device allocation) against lsh_lighting_tiles.

Both entry points are called through the C ABI on the same inputs; every
output must be identical bit for bit, on two calls in a row (the second with
fewer tiles, so a stale tail of the shared static table would show), and a
tile count beyond the static table's 4096 is rejected before any work.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from lunarscout_amd import synthetic

pytestmark = pytest.mark.gpu

PIXELS = 128 * 128


def _outputs(session, n, T, E):
    from lunarscout_amd._lib import DeviceBuffer

    shapes = {"sun_bytes": ((n, T, PIXELS), np.uint8), "sun_fraction": ((n, T, PIXELS), np.float32),
              "sun_mean": ((n, PIXELS), np.float32), "earth_margin": ((n, E, PIXELS), np.float32),
              "earth_visible": ((n, PIXELS), np.float32), "psr": ((n, PIXELS), np.uint8),
              "psr_margin": ((n, PIXELS), np.float32)}
    out = {}
    for name, (shape, dtype) in shapes.items():
        buf = DeviceBuffer.empty(session, shape, dtype)
        buf.zero()
        out[name] = buf
    return out


def _call(entry, session, dem_id, tiles, horizons, vectors, outputs):
    from lunarscout_amd import _lib

    flags = (_lib.LIGHT_SUN_BYTES | _lib.LIGHT_SUN_FRACTION | _lib.LIGHT_SUN_MEAN |
             _lib.LIGHT_EARTH_MARGIN | _lib.LIGHT_EARTH_VISIBLE | _lib.LIGHT_PSR |
             _lib.LIGHT_PSR_MARGIN)
    sun, earth, psr = vectors
    tile_arr = np.ascontiguousarray(np.array(tiles, dtype=np.int32).reshape(-1, 2))
    o = outputs
    return entry(session.handle, dem_id, tile_arr.ctypes.data, len(tiles), horizons.ptr, 1440, 0,
                 sun.ptr, sun.shape[0], earth.ptr, earth.shape[0], psr.ptr, psr.shape[0],
                 0.27, float(np.float32(0.545)), flags, o["sun_bytes"].ptr, o["sun_fraction"].ptr,
                 o["sun_mean"].ptr, o["earth_margin"].ptr, o["earth_visible"].ptr, o["psr"].ptr,
                 o["psr_margin"].ptr)


def test_static_entry_point_matches_the_allocating_one_on_two_calls():
    from lunarscout_amd import _lib
    from lunarscout_amd._lib import DeviceBuffer
    from lunarscout_amd.hip_backend import HipSession
    from lunarscout_amd.lighting import reduce_sun_vectors_for_psr

    s = HipSession(0)
    lib = _lib.load_library()
    dem = synthetic.fbm_crater_dem(512, 20.0, 4)
    dem_id = s.dem_id(dem)
    rng = np.random.default_rng(8)
    sun = synthetic.sun_vectors(dem, 40, 6.0, seed=3)
    earth = synthetic.earth_vectors(30, 6.0, seed=4)
    psr, _ = reduce_sun_vectors_for_psr(dem, sun)
    vectors = [DeviceBuffer.from_host(s, v) for v in (sun, earth, psr)]
    grid = [(x, y) for y in range(0, 512, 128) for x in range(0, 512, 128)]
    for call, count in enumerate((7, 3)):
        tiles = [grid[(i * 5 + call) % len(grid)] for i in range(count)]
        hz = rng.normal(0.0, 2.0, (count, PIXELS, 1440)).astype(np.float32)
        horizons = DeviceBuffer.from_host(s, hz)
        want = _outputs(s, count, len(sun), len(earth))
        got = _outputs(s, count, len(sun), len(earth))
        _lib.check(_call(lib.lsh_lighting_tiles, s, dem_id, tiles, horizons, vectors, want),
                   "lsh_lighting_tiles")
        _lib.check(_call(lib.lsh_lighting_tiles_static, s, dem_id, tiles, horizons, vectors, got),
                   "lsh_lighting_tiles_static")
        s.synchronize()
        for name in want:
            a, b = want[name].to_host(), got[name].to_host()
            assert a.tobytes() == b.tobytes(), (call, name)
        print(f"call {call + 1}: {count} tiles, all 7 outputs identical", flush=True)
        for buf in (*want.values(), *got.values(), horizons):
            buf.free()
    # beyond the static table: rejected, outputs untouched
    tiles = [grid[0]] * 4097
    tiny = DeviceBuffer.empty(s, (1,), np.float32)
    status = _call(lib.lsh_lighting_tiles_static, s, dem_id, tiles, tiny, vectors,
                   {k: tiny for k in ("sun_bytes", "sun_fraction", "sun_mean", "earth_margin",
                                      "earth_visible", "psr", "psr_margin")})
    assert status == _lib.LSH_EINVAL and "4096" in _lib.last_error()
    tiny.free()
    for v in vectors:
        v.free()
    s.close()
