"""The drop-in seams accept the reference's own contract objects.

The reference's pipeline hands ``subpatch_hierarchical_all_passes`` its
``contract.PyramidArrays`` (level0, mips, levels, map_parameters float32[11],
projection_parameters float32[6]; contract.py:274-337, built by
pyramid.py:78-107) and the lighting sessions its ``DemGrid`` whose
``ProjectionParameters`` (geometry.py:24-38) has the six named fields and no
array method. These tests build look-alikes of those types here (the reference
is not imported) from the arrays its captured fixtures hold, the fields
``PyramidArrays.from_artifact`` reads, and pass them through the seams.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import pytest

from tests import fixtures as fx
from lunarscout_amd.errors import ContractValidationError
from lunarscout_amd.geometry import projection_array
from lunarscout_amd.hip_backend import dem_of_pyramid


@dataclass(frozen=True)
class ReferenceProjection:
    """The field layout of the reference's ProjectionParameters (no as_array method)."""

    radius_m: float
    latitude_origin_rad: float
    longitude_origin_rad: float
    scale: float
    false_easting_m: float
    false_northing_m: float


@dataclass(frozen=True)
class ReferenceDemGrid:
    elevation_m: np.ndarray
    geo_transform: np.ndarray
    projection: ReferenceProjection

    @property
    def width(self):
        return self.elevation_m.shape[1]

    @property
    def height(self):
        return self.elevation_m.shape[0]


@dataclass(frozen=True)
class ReferencePyramidArrays:
    level0: np.ndarray
    mips: np.ndarray
    levels: np.ndarray
    map_parameters: np.ndarray
    projection_parameters: np.ndarray


def reference_pyramid(prefix: str) -> ReferencePyramidArrays:
    """What PyramidArrays.from_artifact builds from a fixture's arrays (contract.py:317-337)."""
    a = fx.rays_arrays()
    levels, mips = fx.pyramid_levels(prefix)
    return ReferencePyramidArrays(
        np.ascontiguousarray(a[f"{prefix}__elevation_m"], dtype=np.float32), mips, levels,
        np.ascontiguousarray(a[f"{prefix}__pyramid__map_parameters"], dtype=np.float32),
        np.ascontiguousarray(a[f"{prefix}__pyramid__projection_parameters"], dtype=np.float32))


PREFIXES = ("multi_dem_outer_obstacle_east__dem_0", "multi_dem_outer_obstacle_east__dem_1")


@pytest.mark.parametrize("prefix", PREFIXES)
def test_reference_pyramid_arrays_describe_the_fixture_dem(prefix):
    pyramid = reference_pyramid(prefix)
    dem, mips = dem_of_pyramid(pyramid)
    want = fx.dem(prefix)
    assert (dem.width, dem.height) == (want.width, want.height)
    np.testing.assert_array_equal(dem.elevation_m, want.elevation_m)
    # the float32 map parameters hold the transform the march reads in float32
    np.testing.assert_array_equal(dem.geo_transform.astype(np.float32),
                                  want.geo_transform.astype(np.float32))
    assert np.float32(dem.map_resolution_m) == np.float32(want.map_resolution_m)
    assert np.float32(dem.projection.radius_m) == np.float32(want.projection.radius_m)
    assert mips.tobytes() == pyramid.mips.tobytes()


def test_reference_pyramid_arrays_are_validated():
    pyramid = reference_pyramid(PREFIXES[0])
    bad_levels = ReferencePyramidArrays(pyramid.level0, pyramid.mips, pyramid.levels[:-1],
                                        pyramid.map_parameters, pyramid.projection_parameters)
    with pytest.raises(ContractValidationError, match="level metadata"):
        dem_of_pyramid(bad_levels)
    with pytest.raises(ContractValidationError, match="level0"):
        dem_of_pyramid(object())


def test_projection_fields_are_read_by_name():
    p = ReferenceProjection(1737400.0, -np.pi / 2, 0.25, 1.0, 10.0, -20.0)
    np.testing.assert_array_equal(projection_array(p),
                                  [1737400.0, -np.pi / 2, 0.25, 1.0, 10.0, -20.0])
    np.testing.assert_array_equal(projection_array([1, 2, 3, 4, 5, 6]), [1, 2, 3, 4, 5, 6])
    with pytest.raises(ContractValidationError):
        projection_array([1, 2, 3])


def test_cpu_lighting_seam_takes_a_reference_dem_grid():
    from lunarscout_amd import synthetic
    from lunarscout_amd.geometry import DemGrid, ProjectionParameters
    from lunarscout_amd.lighting import CpuLightingSession

    ours = synthetic.fbm_crater_dem(128, 20.0, 3)
    p = ours.projection
    theirs = ReferenceDemGrid(ours.elevation_m, ours.geo_transform,
                              ReferenceProjection(p.radius_m, p.latitude_origin_rad,
                                                  p.longitude_origin_rad, p.scale,
                                                  p.false_easting_m, p.false_northing_m))
    assert isinstance(ours, DemGrid) and isinstance(ours.projection, ProjectionParameters)
    rng = np.random.default_rng(2)
    horizons = rng.uniform(-2.0, 3.0, (128, 128, 1440)).astype(np.float32)
    sun = synthetic.sun_vectors(ours, 12, 6.0, seed=3)
    cpu = CpuLightingSession(threads=2)
    a = np.stack(list(cpu.iter_patch_tiles(ours, horizons, sun, tile_y=0, tile_x=0)))
    b = np.stack(list(cpu.iter_patch_tiles(theirs, horizons, sun, tile_y=0, tile_x=0)))
    np.testing.assert_array_equal(a, b)
    mask_a = cpu.compute_patch(ours, horizons, sun, tile_y=0, tile_x=0)
    mask_b = cpu.compute_patch(theirs, horizons, sun, tile_y=0, tile_x=0)
    np.testing.assert_array_equal(mask_a, mask_b)
