"""CPU restatement of the reference's typed/coordination active regions and
Gaussian start displacement, in plain loops.

TEST INFRASTRUCTURE ONLY: the checker of akmc_amd/regions.py.  Nothing in
the product path imports it.

Follows optim/SaddleSearch.cpp (constructAtoms styles 4 and 5 :140-349,
makeGaussDisplacement :416-491) and Common/Basis.cpp
(calculateCoordNumber :1300-1360, getDistance :783-807) for orthorhombic
cells.  makeRandomDisplacement and the style-2 sphere are the C oracle's
ako_random_displacement and ako_active_sphere.  Pinned by
tests/test_regions.py: coordination numbers of a perfect lattice and of a
vacancy's shell, Gaussian weights at known distances, the unweighted start
direction, and scripted draws through the rejection loops.
"""
from __future__ import annotations

import math


def _dist2(a, b, lat):
    L = (lat[0], lat[4], lat[8])
    s = 0.0
    for k in range(3):
        d = a[k] - b[k]
        f = d / L[k]
        f -= _rint(f)
        s += (f * L[k]) ** 2
    return s


def _rint(x):
    """C rint: round half to even"""
    r = math.floor(x + 0.5)
    if r - x == 0.5 and r % 2 != 0:
        r -= 1
    return r


def coordination(pos, types, lat, typ1, typ2, dist):
    n = len(types)
    out = [0] * n
    for i in range(n):
        if types[i] != typ1:
            continue
        for j in range(n):
            if types[j] != typ2 or i == j:
                continue
            if _dist2([pos[k][i] for k in range(3)], [pos[k][j] for k in range(3)], lat) <= dist * dist:
                out[i] += 1
    return out


def pick(draws, n, ok):
    """the rejection loop: j = int(u n) for successive draws until ok(j)"""
    for u in draws:
        j = int(u * n)
        if ok(j):
            return j
    raise ValueError("draws exhausted")


def gauss_displacement(pos, frozen, centre, u, d0, width, lat):
    """makeGaussDisplacement with the uniforms u (3 per free atom, atom
    order): returns (displacement per free atom [[dx, dy, dz]...],
    rVector per free atom)"""
    free = [i for i in range(len(frozen)) if frozen[i] == 0]
    d, wgt = [], []
    for a, i in enumerate(free):
        d.append([u[3 * a] - 0.5, u[3 * a + 1] - 0.5, u[3 * a + 2] - 0.5])
        r2 = _dist2([pos[k][i] for k in range(3)], centre, lat)
        wgt.append(math.exp(-r2 / (2.0 * width ** 2)))
    m = len(free)
    mean = [sum(d[a][k] for a in range(m)) / m for k in range(3)]
    for a in range(m):
        for k in range(3):
            d[a][k] -= mean[k]
    norm = math.sqrt(sum(d[a][k] ** 2 for a in range(m) for k in range(3)))
    rvec = [[d[a][k] / norm for k in range(3)] for a in range(m)]
    disp = [[d[a][k] * d0 * wgt[a] for k in range(3)] for a in range(m)]
    return disp, rvec
