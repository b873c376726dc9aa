"""The KMC step bookkeeping: the oracle (oracle/kmc_oracle.py) pinned by
closed forms and brute force, then the library's host functions
(akmc_kmc_* in libakmc_hip.so, no device calls) against the oracle on the
same inputs.  The reference ships no KMC fixtures, so the pins are:
Arrhenius values at chosen energies, exact selection intervals, the
exponential clock's closed forms, single-atom displacements across
periodic images, and recycling cases worked out by hand."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import kmc_oracle as K

LAT = np.diag([10.0, 12.0, 14.0]).reshape(9)


def _lib():
    from akmc_amd import lib
    return lib.load()


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int))


# ------------------------------------------------------------------ oracle pins

def test_rate_closed_forms():
    T, nu = 1000.0, 1.0e13
    assert K.rate(0.0, nu, T) == nu
    # Ea = kB T ln 2 halves the rate; Ea = kB T gives nu / e
    assert K.rate(K.BOLTZMANN * T * math.log(2.0), nu, T) == pytest.approx(nu / 2, rel=1e-14)
    assert K.rate(K.BOLTZMANN * T, nu, T) == pytest.approx(nu / math.e, rel=1e-14)
    # a 0.5 eV barrier at 300 K: exp(-0.5 / 0.025852) = 3.99e-9 (CODATA kB)
    assert K.rate(0.5, 1.0, 300.0) == pytest.approx(math.exp(-0.5 / (8.617333262e-5 * 300.0)), rel=1e-14)
    # the gdc example's window (docs/examples/gdc/control: T 1000, mincap 0.1, window 50)
    assert K.max_cap(0.1, 50.0, 1000.0) == pytest.approx(0.1 + 50 * 1000 * 8.617333262e-5, rel=1e-15)


def test_selection_intervals_are_exact():
    rates = [1.0, 3.0, 0.0, 4.0]          # cumulative 1, 4, 4, 8 of total 8
    assert K.select_event(rates, 0.0) == (0, 8.0)
    assert K.select_event(rates, 0.124999)[0] == 0
    assert K.select_event(rates, 0.125)[0] == 1   # u R = 1 is not > 1: the next event
    assert K.select_event(rates, 0.4999)[0] == 1
    assert K.select_event(rates, 0.5)[0] == 3     # the zero-rate event is never chosen
    assert K.select_event(rates, 0.99999)[0] == 3
    assert K.select_event([], 0.3) == (-1, 0.0)
    # frequencies over a uniform grid of u equal r_i / R
    u = (np.arange(80000) + 0.5) / 80000
    counts = np.bincount([K.select_event(rates, x)[0] for x in u], minlength=4)
    assert np.allclose(counts / u.size, np.array(rates) / 8.0)


def test_clock_closed_forms():
    R = 2.5e6
    assert K.time_increase(R, math.exp(-1.0)) == pytest.approx(1.0 / R, rel=1e-15)
    assert K.time_increase(R, 1.0) == 0.0
    assert K.time_increase(0.0, 0.5) == 0.0
    # mean of -ln(u) / R over a midpoint grid of u -> 1 / R
    u = (np.arange(200000) + 0.5) / 200000
    mean = np.mean([K.time_increase(R, x) for x in u[::50]])
    assert mean == pytest.approx(1.0 / R, rel=2e-3)


def _pair(n=6, seed=0):
    rng = np.random.default_rng(seed)
    b1 = rng.random((3, n)) * np.array([[10.0], [12.0], [14.0]])
    return b1, b1.copy()


def test_check_basin_single_atom_and_images():
    b1, b2 = _pair()
    assert not K.check_basin(b1, b2, LAT, 0.5, 3.0)
    b2[:, 2] += [0.49, 0.0, 0.0]
    assert not K.check_basin(b1, b2, LAT, 0.5, 3.0)
    b2[:, 2] = b1[:, 2] + [0.5, 0.0, 0.0]           # 0.5 (to rounding of the sum): counts (>=)
    assert K.check_basin(b1, b2, LAT, 0.4999999, 3.0)
    assert not K.check_basin(b1, b2, LAT, 0.5000001, 3.0)
    # a whole lattice vector is no displacement; 10 - 0.6 is 0.6 by minimum image
    b2[:, 2] = b1[:, 2] + [10.0, -12.0, 14.0]
    assert not K.check_basin(b1, b2, LAT, 0.5, 3.0)
    b2[:, 2] = b1[:, 2] + [9.4, 0.0, 0.0]
    assert K.check_basin(b1, b2, LAT, 0.5, 3.0)
    # one atom beyond max_radius vetoes, whatever the others do
    b2[:, 4] = b1[:, 4] + [0.0, 0.0, 3.01]
    assert not K.check_basin(b1, b2, LAT, 0.5, 3.0)
    b2[:, 4] = b1[:, 4] + [0.0, 0.0, 3.0]
    assert K.check_basin(b1, b2, LAT, 0.5, 3.0)


def test_basin2_kick_closed_form():
    b1, sad = _pair(5, 1)
    d = np.zeros((3, 5))
    d[:, 1] = [0.8, -0.3, 0.2]
    d[:, 3] = [0.0, 0.1, -0.1]
    sad = b1 + d
    sad[0, 1] -= 10.0                  # the saddle atom sits one image away
    frozen = np.array([0, 0, 0, 1, 0])
    out = K.basin2_start(sad, b1, frozen, LAT, 0.2)
    norm = 0.2 * np.sqrt((d * d).sum())
    exp = sad.copy()
    exp[:, 1] += norm * d[:, 1]
    assert np.allclose(out, exp, atol=1e-13, rtol=0)
    assert np.array_equal(out[:, 3], sad[:, 3])    # frozen atoms stay at the saddle


def test_kick_pattern_differs_from_check_pattern_only_off_orthogonal():
    """The kick and checkBasin read the reciprocal vectors with different
    index patterns; on orthogonal cells both are the minimum image."""
    tri = np.array([10.0, 0, 0, 2.0, 9.0, 0, 1.0, 1.5, 8.0])
    b1 = np.zeros((3, 1))
    s = np.array([[1.3], [-0.7], [0.9]])
    k_orth = K.basin2_start(s, b1, [0], LAT, 1.0) - s
    assert np.allclose(k_orth, np.sqrt((s * s).sum()) * s)
    k_tri = K.basin2_start(s, b1, [0], tri, 1.0) - s
    assert not np.allclose(k_tri, np.sqrt((s * s).sum()) * s)


def test_recycling_by_hand():
    lat = np.diag([8.0, 16.0, 16.0]).reshape(9)   # power-of-two edges: exact image arithmetic
    cen = [np.array([1.0, 1.0, 1.0]), np.array([4.0, 1.0, 1.0]), np.array([7.5, 1.0, 1.0]),
           np.array([1.0, 7.0, 1.0])]
    # chosen 0; distances 3, 1.5 (image), 6
    assert K.recycle(cen, 0, lat, 2.0, False, 1, 100.0) == []
    assert K.recycle(cen, 0, lat, 2.0, True, 1, 100.0) == [1, 3]
    assert K.recycle(cen, 0, lat, 1.5, True, 1, 100.0) == [1, 2, 3]     # r >= radius
    assert K.recycle(cen, 0, lat, 2.0, True, 2, 50.0, u=[0, 0.2, 0.9, 0.6]) == [1]
    # superBasin: the reference compares the excluded list's entries (counters)
    # with the event's displaced atom numbers
    disp = [[], [5], [], [2]]
    assert K.recycle(cen, 0, lat, 2.0, True, 1, 100.0, displaced=disp, excluded=[0, 0, 2, 0, 0, 0],
                     super_basin=True) == [1]


def test_excluded_list_history():
    assert K.update_excluded([0, 3, 1, 0], [0, 2], 5) == [5, 2, 5, 0]


# ------------------------------------------------------ library vs oracle

def test_library_rates_selection_clock_match_oracle():
    L = _lib()
    rng = np.random.default_rng(4)
    for _ in range(50):
        ea = float(rng.uniform(0.0, 2.0))
        T = float(rng.uniform(300, 1500))
        assert L.akmc_kmc_rate(ea, 1e13, T) == K.rate(ea, 1e13, T)
    assert L.akmc_kmc_max_cap(0.1, 50.0, 1000.0) == K.max_cap(0.1, 50.0, 1000.0)
    for n in (0, 1, 3, 17):
        rates, rp = _d(rng.random(n) * 10.0 ** rng.integers(-3, 3, n))
        for u in list(rng.random(20)) + [0.0]:
            tot = C.c_double()
            i = L.akmc_kmc_select(n, rp, float(u), C.byref(tot))
            ei, et = K.select_event(list(rates), float(u))
            assert i == ei and tot.value == et
    for R in (0.0, 1e-20, 3.0, 1e9):
        for u in (0.1, 0.5, 0.999):
            assert L.akmc_kmc_time_increase(R, u) == K.time_increase(R, u)


def test_library_basin_checks_match_oracle():
    L = _lib()
    rng = np.random.default_rng(8)
    lat, lp = _d(LAT)
    for trial in range(40):
        n = int(rng.integers(1, 30))
        b1 = rng.random((3, n)) * np.array([[10.0], [12.0], [14.0]])
        b2 = b1 + rng.normal(0.0, 0.4, (3, n)) + np.array([[10.0], [0.0], [-14.0]]) * rng.integers(-1, 2, (1, n))
        radius, rmax = 0.5, float(rng.uniform(0.8, 2.0))
        a1, p1 = _d(b1)
        a2, p2 = _d(b2)
        got = L.akmc_kmc_check_basin(n, p1, p2, lp, radius, rmax)
        assert got == int(K.check_basin(b1, b2, LAT, radius, rmax)), trial
        atoms, ap = _i(np.zeros(n))
        m = L.akmc_kmc_displaced_atoms(n, p2, p1, lp, radius, ap)
        assert list(atoms[:m]) == K.displaced_atoms(b2, b1, LAT, radius)


def test_library_recycling_matches_oracle():
    L = _lib()
    rng = np.random.default_rng(12)
    lat, lp = _d(LAT)
    for trial in range(60):
        n = int(rng.integers(1, 12))
        cen = rng.random((n, 3)) * [10.0, 12.0, 14.0]
        chosen = int(rng.integers(0, n))
        mode = int(rng.integers(1, 3))
        u = rng.random(n)
        sb = bool(rng.integers(0, 2))
        excluded = list(rng.integers(0, 4, 8))
        disp = [list(rng.choice(8, size=int(rng.integers(0, 3)), replace=False)) for _ in range(n)]
        blocked = [int(any(ex == a for ex in excluded for a in disp[i])) if sb else 0 for i in range(n)]
        c, cp = _d(cen.reshape(-1))
        uu, up = _d(u)
        bl, bp = _i(blocked)
        keep, kp = _i(np.zeros(n))
        got = L.akmc_kmc_recycle(n, cp, chosen, lp, 3.0, 1, mode, 40.0, up, bp, kp)
        exp = K.recycle(list(cen), chosen, LAT, 3.0, True, mode, 40.0, u=u, displaced=disp, excluded=excluded,
                        super_basin=sb)
        assert list(np.nonzero(keep)[0]) == exp and got == len(exp), trial
    ex, ep = _i([0, 3, 1, 0])
    d, dp = _i([0, 2])
    assert L.akmc_kmc_update_excluded(4, ep, 2, dp, 5) == 0
    assert list(ex) == K.update_excluded([0, 3, 1, 0], [0, 2], 5)
