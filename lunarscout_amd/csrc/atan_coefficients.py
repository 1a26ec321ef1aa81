"""Coefficients of the float64 arctangent in lsh_lighting_core.h.

atan(q) = q + q^3 P(q^2) on |q| <= tan(pi/8), the range left after the
reductions atan(y/x) = pi/2 - atan(x/y) and atan(t) = pi/4 + atan((t-1)/(t+1)).
P is the Chebyshev interpolant of g(z) = (atan(sqrt z) - sqrt z) / z^(3/2) at
n + 1 Chebyshev points of [0, tan^2(pi/8)] (computed with 60-digit arithmetic)
rewritten in powers of z and rounded to float64; the script reports the worst
error of q + q^3 P(q^2) against atan on a dense grid of the range.

    python lunarscout_amd/csrc/atan_coefficients.py [degree]
"""

from __future__ import annotations

import sys

import mpmath as mp

mp.mp.dps = 60


def coefficients(degree: int) -> list[float]:
    zmax = mp.tan(mp.pi / 8) ** 2

    def g(z):
        if z == 0:
            return mp.mpf(-1) / 3
        r = mp.sqrt(z)
        return (mp.atan(r) - r) / (z * r)

    n = degree + 1
    nodes = [zmax / 2 * (1 + mp.cos(mp.pi * (k + mp.mpf(1) / 2) / n)) for k in range(n)]
    values = [g(z) for z in nodes]
    # Vandermonde solve in 60 digits: the interpolant in powers of z
    matrix = mp.matrix([[z ** j for j in range(n)] for z in nodes])
    return [float(c) for c in mp.lu_solve(matrix, mp.matrix(values))]


def worst_error(c: list[float], samples: int = 20001) -> float:
    qmax = float(mp.tan(mp.pi / 8))
    worst = mp.mpf(0)
    for i in range(samples):
        q = mp.mpf(qmax) * i / (samples - 1)
        z = q * q
        p = mp.mpf(c[-1])
        for coef in reversed(c[:-1]):
            p = p * z + coef
        worst = max(worst, abs(q + q * z * p - mp.atan(q)))
    return float(worst)


if __name__ == "__main__":
    degree = int(sys.argv[1]) if len(sys.argv) > 1 else 10
    c = coefficients(degree)
    for value in c:
        print(f"    {value!r},")
    print(f"# degree {degree}: worst |error| {worst_error(c):.3e} on [0, tan(pi/8)]")
