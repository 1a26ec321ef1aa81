"""Write golden vectors from the reference's pure-numpy functions that no
captured fixture pins.

Run in the build container only (it reads the reference sources under
/root/reference and imports three of its pure-numpy modules -- contract.py,
geometry.py, psr.py, with file_format.py for two constants -- by path, without
the package's own __init__):

    python tests/golden/make_reference_vectors.py

It writes tests/golden/reference_polar.npz (inputs are regenerated by the tests
from the seeds recorded in the file, with lunarscout_amd.synthetic):

* ``segments``: float32 [32, 36, 2, 18], the reference's pure-numpy segment
  builder (``build_subpatch_segments`` over ``SubpatchSegmentCache``,
  geometry.py:490-573) for the 32-px tile at (96, 64) of a two-DEM
  polar-stereographic site at 89.5 S (primary 256^2 @ 20 m seed 21, regional
  256^2 @ 120 m seed 22), 32 azimuths, 8-px subpatches, observer 0 m;
* ``psr_indices``: int64, the five-viewpoint PSR Sun-vector reduction
  (``reduce_sun_vectors_for_psr``, psr.py:61-91) of 4380 synthetic Sun vectors
  (2-hour steps, seed 23) over the primary DEM;
* ``sun_fraction``: float32 [3000], ``_sun_fraction_reference`` (lightmap.py:32-73)
  for 30 random 1440-bin horizon rows x 100 random (azimuth, elevation) pairs
  (``fraction_cases``, seed 24);
* ``cbin_payload``: uint8, the ``.cbin`` payload of ``_encode_horizons_python``
  (file_format.py:39-73) for 16 random 1440-bin rows with jumps, clamped
  values and half-step quantisation ties (``codec_input``, seed 25): per pixel a
  little-endian u16 length and the block.
"""

from __future__ import annotations

import importlib
import sys
import types
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
REFERENCE = Path("/root/reference/src/lunarscout/_numba_horizon")

sys.path.insert(0, str(REPO))

from lunarscout_amd import synthetic  # noqa: E402

SEEDS = {"primary": 21, "regional": 22, "sun": 23}
TILE = (96, 64)  # (column, row)
TILE_WIDTH = 32
AZIMUTHS = 32


def site():
    primary = synthetic.fbm_crater_dem(256, 20.0, SEEDS["primary"])
    regional = synthetic.fbm_crater_dem(256, 120.0, SEEDS["regional"])
    sun = synthetic.sun_vectors(primary, 4380, 2.0, seed=SEEDS["sun"])
    return primary, regional, sun


def fraction_cases():
    """30 horizon rows [30, 1440] and 3000 (row, azimuth, elevation) cases."""
    rng = np.random.default_rng(24)
    az = np.arange(1440) * 0.25
    rows = np.stack([rng.uniform(-1, 2) + rng.uniform(0.2, 3) * np.sin(np.radians(az) * k +
                     rng.uniform(0, 6.3)) + rng.normal(0, 0.3, 1440)
                     for k in rng.integers(1, 9, 30)]).astype(np.float32)
    which = np.repeat(np.arange(30), 100)
    azimuth = rng.uniform(0.0, 360.0, 3000)
    azimuth[:20] = [0.0, 0.1, 0.27, 0.395, 0.4, 359.9, 359.75, 90.0, 180.0, 270.0,
                    0.125, 0.5, 1.0, 359.0, 0.3950001, 0.39499998, 45.25, 0.0001, 359.9999, 120.37]
    elevation = rng.uniform(-1.5, 4.5, 3000)
    return rows, which, azimuth, elevation


def codec_input():
    """float32 degrees [16, 1440] exercising every branch of the encoder."""
    rng = np.random.default_rng(25)
    x = np.cumsum(rng.normal(0.0, 0.02, (16, 1440)), axis=1) + rng.uniform(-10, 10, (16, 1))
    x[:, 100:110] += 30.0                      # two-byte deltas
    x[0, 200:260] = 75.0                       # clamped above +50
    x[1, 300:380] = -80.0                      # clamped below -50
    x[2, 500] = -49.9; x[2, 501] = 49.9        # delta beyond the 15-bit clamp
    q = 50.0 / 32767.0
    x[3, :64] = (np.arange(64) - 32 + 0.5) * q  # quantisation ties
    x[4, :] = np.where(np.arange(1440) % 2, 0.0, 62 * q)  # 1-byte edge 62 / 63
    return x.astype(np.float32)


def _reference_modules():
    package = types.ModuleType("reference_horizon")
    package.__path__ = [str(REFERENCE)]
    sys.modules["reference_horizon"] = package
    return (importlib.import_module("reference_horizon.geometry"),
            importlib.import_module("reference_horizon.psr"),
            importlib.import_module("reference_horizon.lightmap"),
            importlib.import_module("reference_horizon.file_format"))


def main() -> None:
    geometry, psr, lightmap, file_format = _reference_modules()
    primary, regional, sun = site()

    def reference_dem(dem):
        p = dem.projection
        return geometry.DemGrid(
            np.ascontiguousarray(dem.elevation_m, dtype=np.float32),
            np.ascontiguousarray(dem.geo_transform, dtype="<f8"),
            geometry.ProjectionParameters(p.radius_m, p.latitude_origin_rad,
                                          p.longitude_origin_rad, p.scale, p.false_easting_m,
                                          p.false_northing_m))

    dems = [reference_dem(primary), reference_dem(regional)]
    segments, _, _ = geometry.build_subpatch_segments(
        dems, tile_column=TILE[0], tile_row=TILE[1], tile_width=TILE_WIDTH,
        azimuth_count=AZIMUTHS, maximum_distance_m=1_000_000.0, observer_elevation_m=0.0,
        subpatch_size=8, grid_convergence=geometry.GridConvergenceInput(0.0, 0.0, 0.0))
    _, indices = psr.reduce_sun_vectors_for_psr(dems[0], sun)
    rows, which, azimuth, elevation = fraction_cases()
    fractions = np.array([lightmap._sun_fraction_reference(rows[r], float(a), float(e))
                          for r, a, e in zip(which, azimuth, elevation)], dtype=np.float32)
    encoded, lengths = file_format._encode_horizons_python(codec_input())
    payload = b"".join(int(n).to_bytes(2, "little") + encoded[i, :n].tobytes()
                       for i, n in enumerate(lengths))
    out = HERE / "reference_polar.npz"
    np.savez_compressed(out, segments=segments.astype(np.float32), psr_indices=indices,
                        sun_fraction=fractions,
                        cbin_payload=np.frombuffer(payload, dtype=np.uint8),
                        tile=np.array(TILE, np.int32), tile_width=np.int32(TILE_WIDTH),
                        azimuths=np.int32(AZIMUTHS),
                        seeds=np.array([SEEDS["primary"], SEEDS["regional"], SEEDS["sun"]],
                                       np.int32))
    print(f"wrote {out}: segments {segments.shape}, {len(indices)} PSR vectors kept, "
          f"{len(fractions)} sun fractions, {len(payload)} payload bytes")


if __name__ == "__main__":
    main()
