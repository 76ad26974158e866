"""Derives the arctangent polynomial used by the pair kernels.

The pair kernels need atan(t) only for t = tan(a/2) in [0, 1] (a is the S3 angle to
the nearer of +-y, so a <= pi/2).  There atan(t) = t * P(t^2) with P a polynomial,
and no range reduction is needed.  P is the minimax fit of relative error, found by
the Remez exchange algorithm (E. Ya. Remez, 1934) on g(z) = atan(sqrt z)/sqrt z over
z in [0, 1] with weight 1/g.  Run this file to print the coefficients that
``qvm_device.h`` holds (highest degree first) and the fit's bound.
"""
import numpy as np


def g(z):
    z = np.asarray(z, dtype=np.float64)
    out = np.ones_like(z)
    nz = z > 0
    r = np.sqrt(z[nz])
    out[nz] = np.arctan(r) / r
    return out


def remez(degree, iterations=30):
    m = degree + 2
    # Chebyshev extrema as the first reference set.
    k = np.arange(m)
    nodes = 0.5 - 0.5 * np.cos(np.pi * k / (m - 1))
    grid = np.linspace(0.0, 1.0, 200001)
    coef = None
    for _ in range(iterations):
        a = np.zeros((m, m))
        for i, z in enumerate(nodes):
            a[i, :degree + 1] = z ** np.arange(degree + 1)
            a[i, degree + 1] = (-1) ** i * g(np.array([z]))[0]
        sol = np.linalg.solve(a, g(nodes))
        coef = sol[:degree + 1]
        err = (np.polyval(coef[::-1], grid) - g(grid)) / g(grid)
        # New reference set: the extremum of each sign run.
        sign = np.sign(err)
        sign[sign == 0] = 1
        cuts = np.flatnonzero(np.diff(sign)) + 1
        runs = np.split(np.arange(len(grid)), cuts)
        ext = np.array([r[np.argmax(np.abs(err[r]))] for r in runs])
        if len(ext) < m:
            break
        # Keep m alternating extrema with the largest errors.
        while len(ext) > m:
            if abs(err[ext[0]]) < abs(err[ext[-1]]):
                ext = ext[1:]
            else:
                ext = ext[:-1]
        nodes = grid[ext]
    err = (np.polyval(coef[::-1], grid) - g(grid)) / g(grid)
    return coef, float(np.max(np.abs(err)))


if __name__ == "__main__":
    for degree in (6, 7, 8):
        coef, bound = remez(degree)
        print(f"degree {degree}: max relative error {bound:.3e}")
        for c in coef[::-1]:
            print(f"    {np.float32(c)!r}  # {c:.17e}")
