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
  (2-hour steps, seed 23) over the primary DEM.
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


def _reference_modules():
    package = types.ModuleType("reference_horizon")
    package.__path__ = [str(REFERENCE)]
    sys.modules["reference_horizon"] = package
    return (importlib.import_module("reference_horizon.geometry"),
            importlib.import_module("reference_horizon.psr"))


def main() -> None:
    geometry, psr = _reference_modules()
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
    out = HERE / "reference_polar.npz"
    np.savez_compressed(out, segments=segments.astype(np.float32), psr_indices=indices,
                        tile=np.array(TILE, np.int32), tile_width=np.int32(TILE_WIDTH),
                        azimuths=np.int32(AZIMUTHS),
                        seeds=np.array([SEEDS["primary"], SEEDS["regional"], SEEDS["sun"]],
                                       np.int32))
    print(f"wrote {out}: segments {segments.shape}, {len(indices)} PSR vectors kept")


if __name__ == "__main__":
    main()
