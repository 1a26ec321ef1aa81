"""Extract the reference's golden data into compact fixtures for this repository.

Reads the data files the reference's own tests hold (captured from the C#/ILGPU
generator) and writes the arrays and metadata the tests here use:

* ``reference_rays.npz`` / ``reference_rays.json``  from
  tests/data/numba_horizon/phase1_reference_rays.{npz,json}
* ``phase4d.json``   from tests/data/numba_horizon/phase4d_production_segments.json
* ``lightmap_csharp.json`` from tests/data/numba_horizon/phase6b_lightmap_csharp.json
* ``psr_csharp.json`` from tests/data/numba_horizon/phase6b_psr_csharp.json

Only data is copied (inputs and expected outputs); nothing from the reference's
source is read or run. Usage: ``python tests/golden/make_fixtures.py``.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

SOURCE = Path("/root/reference/tests/data/numba_horizon")
HERE = Path(__file__).resolve().parent

# Array name suffixes the tests read.
KEEP_SUFFIXES = (
    "__elevation_m", "__geo_transform", "__level_metadata", "__map_parameters",
    "__projection_parameters", "__segment_values", "__centers", "__grid_convergence",
    "__nominal_direction_moon_centered", "__observer_vector_moon_centered_m",
    "__sample_distance_m", "__sample_pixel_x", "__sample_pixel_y", "__sample_latitude_rad",
    "__sample_longitude_rad", "__sample_row", "__sample_column", "__sample_terrain_height_m",
    "__final_degrees", "__final_slopes", "__per_dem_slopes", "__traversal_trace",
)


def _keep(name: str) -> bool:
    if "__pyramid_fixture__" in name:
        return True
    if "__pyramid__level_" in name and not name.endswith("level_metadata") and not name.endswith(
        "level_cell_sizes"
    ):
        # every mip level, except level 0 (identical to the elevation array)
        return not name.endswith("__level_0")
    return name.endswith(KEEP_SUFFIXES)


def main() -> None:
    metadata = json.loads((SOURCE / "phase1_reference_rays.json").read_text(encoding="utf-8"))
    arrays = {}
    with np.load(SOURCE / "phase1_reference_rays.npz", allow_pickle=False) as archive:
        for name in archive.files:
            if _keep(name):
                arrays[name] = np.ascontiguousarray(archive[name])
    np.savez_compressed(HERE / "reference_rays.npz", **arrays)
    slim = {
        key: metadata[key]
        for key in ("conventions", "cases", "pyramid_fixtures", "subpatch_fixtures",
                    "horizon_buffer_fixtures")
    }
    (HERE / "reference_rays.json").write_text(json.dumps(slim), encoding="utf-8")

    phase4d = json.loads(
        (SOURCE / "phase4d_production_segments.json").read_text(encoding="utf-8")
    )
    (HERE / "phase4d.json").write_text(json.dumps(phase4d), encoding="utf-8")
    for source, target in (
        ("phase6b_lightmap_csharp.json", "lightmap_csharp.json"),
        ("phase6b_psr_csharp.json", "psr_csharp.json"),
    ):
        payload = json.loads((SOURCE / source).read_text(encoding="utf-8"))
        (HERE / target).write_text(json.dumps(payload), encoding="utf-8")
    print("wrote", sorted(p.name for p in HERE.iterdir() if p.suffix in (".npz", ".json")))


if __name__ == "__main__":
    main()
