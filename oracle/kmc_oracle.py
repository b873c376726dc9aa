"""CPU restatement of the reference's kinetic Monte Carlo step bookkeeping.

TEST INFRASTRUCTURE ONLY: this module is the checker the tests compare the
library's KMC functions and the device basin-2 stage against.  Nothing in
the product path imports it.

Every function restates one piece of kmc/KineticMonteCarlo.cpp,
kmc/Event.cpp or Common/Basis.cpp in plain Python loops, in the
reference's own index patterns (positions are (3, n) arrays, lattices the
reference's latVector[9] with the lattice vectors as rows, reciprocal
vectors in the layout of Basis::setRcpVec).  The Boltzmann constant lives
in the reference's missing Constants.h; this restatement uses the CODATA
2018 value.  Pinned by tests/test_kmc_oracle.py: Arrhenius closed forms,
exact selection intervals, the exponential clock's closed forms, single-
atom displacements across periodic images and brute-force recycling cases.
"""
from __future__ import annotations

import math

import numpy as np

BOLTZMANN = 8.617333262e-5  # eV / K


def rcp_vector(lat):
    """Basis::setRcpVec (Common/Basis.cpp:987-1010), its storage layout
    included (entries 5 and 7 hold each other's cofactor)."""
    L = [float(v) for v in lat]
    r = [0.0] * 9
    r[0] = L[4] * L[8] - L[7] * L[5]
    r[1] = L[6] * L[5] - L[3] * L[8]
    r[2] = L[3] * L[7] - L[6] * L[4]
    r[3] = L[7] * L[2] - L[1] * L[8]
    r[4] = L[0] * L[8] - L[6] * L[2]
    r[7] = L[6] * L[1] - L[0] * L[7]
    r[6] = L[1] * L[5] - L[4] * L[2]
    r[5] = L[2] * L[3] - L[5] * L[0]
    r[8] = L[0] * L[4] - L[3] * L[1]
    det = L[0] * r[0] + L[1] * r[3] + L[2] * r[6]
    f = 1.0 / det if abs(det) > 0.0 else 1.0
    return [v * f for v in r]


def _wrap_basis(rx, ry, rz, lat, rcp):
    """the minimum-image pattern of Basis::checkBasin and
    Event::getDisplacedAtoms (Common/Basis.cpp:753-763, kmc/Event.cpp:322-332)"""
    xx = rx * rcp[0] + ry * rcp[1] + rz * rcp[2]
    yy = rx * rcp[3] + ry * rcp[4] + rz * rcp[5]
    zz = rx * rcp[6] + ry * rcp[7] + rz * rcp[8]
    xx -= round_half_even(xx)
    yy -= round_half_even(yy)
    zz -= round_half_even(zz)
    return (xx * lat[0] + yy * lat[3] + zz * lat[6],
            xx * lat[1] + yy * lat[4] + zz * lat[7],
            xx * lat[2] + yy * lat[5] + zz * lat[8])


def round_half_even(v):
    """C rint in the default rounding mode"""
    return float(np.rint(v))


def rate(ea, prefactor, temperature):
    """Event rate (kmc/Event.cpp:71,103)"""
    beta = 1.0 / (BOLTZMANN * temperature)
    return prefactor * math.exp(-ea * beta)


def max_cap(min_cap, window, temperature):
    """kmc/kmc.cpp:160"""
    return min_cap + window * temperature * BOLTZMANN


def select_event(rates, u):
    """The event choice of runKMC (kmc/KineticMonteCarlo.cpp:144-169):
    total rate, ranNumber = u * total, first event whose running sum
    exceeds it.  Returns (index or -1, total)."""
    total = 0.0
    for r in rates:
        total += r
    ran = u * total
    acc = 0.0
    for i, r in enumerate(rates):
        acc += r
        if acc > ran:
            return i, total
    return -1, total


def time_increase(total_rate, u):
    """calculateTime (kmc/KineticMonteCarlo.cpp:1151-1161) negated as at :173"""
    dt = 0.0
    if abs(total_rate) > 1.0e-18:
        dt = (1.0 / total_rate) * math.log(u)
    return -1.0 * dt


def basin2_start(saddle, basin1, frozen, lat, basin2_delta):
    """The kick of checkForNewBasin (kmc/KineticMonteCarlo.cpp:1648-1683) in
    its own index pattern (which differs from checkBasin's for non-
    orthogonal cells): Delta = min-image(saddle - basin1) for every atom,
    norm = |Delta| basin2Delta, free atoms start at saddle + norm Delta."""
    rcp = rcp_vector(lat)
    n = saddle.shape[1]
    d = np.zeros((3, n))
    for i in range(n):
        rx = saddle[0, i] - basin1[0, i]
        ry = saddle[1, i] - basin1[1, i]
        rz = saddle[2, i] - basin1[2, i]
        xx = rx * rcp[0] + ry * rcp[1] + rz * rcp[2]
        yy = rx * rcp[3] + ry * rcp[4] + rz * rcp[7]
        zz = rx * rcp[6] + ry * rcp[5] + rz * rcp[8]
        xx -= round_half_even(xx)
        yy -= round_half_even(yy)
        zz -= round_half_even(zz)
        d[0, i] = xx * lat[0] + yy * lat[1] + zz * lat[2]
        d[1, i] = xx * lat[3] + yy * lat[4] + zz * lat[7]
        d[2, i] = xx * lat[6] + yy * lat[5] + zz * lat[8]
    norm = 0.0
    for i in range(n):
        norm += d[0, i] * d[0, i] + d[1, i] * d[1, i] + d[2, i] * d[2, i]
    norm = math.sqrt(norm) * basin2_delta
    out = np.array(saddle, dtype=np.float64, copy=True)
    for i in range(n):
        if frozen[i] == 0:
            for c in range(3):
                out[c, i] = saddle[c, i] + norm * d[c, i]
    return out


def check_basin(basin1, basin2, lat, radius, max_radius):
    """Basis::checkBasin (Common/Basis.cpp:731-781) with globalNo the
    identity: True when some atom moved by >= radius, False as soon as one
    moved by > max_radius."""
    rcp = rcp_vector(lat)
    r2, m2 = radius * radius, max_radius * max_radius
    check = False
    for i in range(basin1.shape[1]):
        rx, ry, rz = _wrap_basis(basin1[0, i] - basin2[0, i], basin1[1, i] - basin2[1, i],
                                 basin1[2, i] - basin2[2, i], lat, rcp)
        rsq = rx * rx + ry * ry + rz * rz
        if rsq >= r2:
            check = True
        if rsq > m2:
            check = False
            break
    return check


def displaced_atoms(pos, basin, lat, radius):
    """Event::getDisplacedAtoms (kmc/Event.cpp:303-341): atoms displaced by
    more than radius."""
    rcp = rcp_vector(lat)
    r2 = radius * radius
    out = []
    for i in range(pos.shape[1]):
        rx, ry, rz = _wrap_basis(pos[0, i] - basin[0, i], pos[1, i] - basin[1, i], pos[2, i] - basin[2, i],
                                 lat, rcp)
        if rx * rx + ry * ry + rz * rz > r2:
            out.append(i)
    return out


def recycle(centres, chosen, lat, radius, recycle_on, mode, percent, u=None, displaced=None, excluded=None,
            super_basin=False):
    """clearEventList (kmc/KineticMonteCarlo.cpp:1163-1265): the indices of
    the events kept.  displaced[i] lists event i's displaced atoms (for the
    superBasin test against the excluded list); u[i] is the draw for
    event i in mode 2."""
    if not recycle_on:
        return []
    rcp = rcp_vector(lat)
    ratio = percent / 100 if mode == 2 else 1.2
    choice = 0.0
    found = False
    ac = centres[chosen]
    kept = []
    for i in range(len(centres)):
        if i == chosen:
            continue
        bc = centres[i]
        rx, ry, rz = _wrap_basis(ac[0] - bc[0], ac[1] - bc[1], ac[2] - bc[2], lat, rcp)
        r = math.sqrt(rx * rx + ry * ry + rz * rz)
        if r >= radius:
            if super_basin:
                found = False
                for ex in excluded:
                    for atm in displaced[i]:
                        if ex == atm:
                            found = True
            if mode == 2:
                choice = u[i]
            if choice < ratio and not found:
                kept.append(i)
    return kept


def update_excluded(excluded, displaced, history):
    """updateExcludedList (kmc/KineticMonteCarlo.cpp:1291-1313)"""
    out = list(excluded)
    for i in range(len(out)):
        if out[i] > 0:
            out[i] -= 1
    for j in displaced:
        out[j] += history
    return out
