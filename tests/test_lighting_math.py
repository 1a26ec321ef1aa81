"""The lighting path's float64 arctangent (lunarscout_amd/csrc/lsh_math.h).

The pixel frame's azimuth and elevation (psr.py:50-58) use a reduced-range
polynomial arctangent instead of the general library routine. Pinned here on
two million pixel-to-body-like vectors (magnitudes 1e-3 .. 1e11 m, every
quadrant, both axes) against an 80-bit reference: at most 3 ulp, i.e. below
1e-15 rad, far under the float32 rounding (3e-5 deg at 360 deg) of the
azimuth the sun-disk walk consumes.
"""

from __future__ import annotations

import numpy as np

from tests import kernel_cpu


def test_arctangent_is_within_three_ulp_everywhere():
    rng = np.random.default_rng(0)
    n = 2_000_000
    mag = 10.0 ** rng.uniform(-3, 11, n)
    ang = rng.uniform(-np.pi, np.pi, n)
    y, x = mag * np.sin(ang), mag * np.cos(ang)
    y[:1000] = 0.0
    x[1000:2000] = 0.0
    x[2000:3000] = -x[2000:3000] ** 2 / mag[2000:3000]  # near the negative axis
    got = kernel_cpu.atan2(y, x)
    ref = np.arctan2(y.astype(np.longdouble), x.astype(np.longdouble))
    ulp = np.spacing(np.abs(ref).astype(np.float64))
    err = np.abs(got.astype(np.longdouble) - ref) / np.maximum(ulp, 1e-300)
    assert float(err.max()) <= 3.0, float(err.max())
    # the quadrant conventions of atan2: (0, -1) is pi, (-1, 0) is -pi/2
    assert kernel_cpu.atan2([0.0, -1.0, 1.0, 0.0], [-1.0, 0.0, 0.0, 1.0]).tolist() == \
        [np.pi, -np.pi / 2, np.pi / 2, 0.0]
