"""Active regions and start displacements (akmc_amd/regions.py) against
the oracle (the C oracle's makeRandomDisplacement and style-2 sphere,
oracle/regions_oracle.py for the rest) on the same draws, with the oracle
pinned by closed forms: coordination numbers of a perfect lattice and of a
vacancy's first shell, Gaussian weights at known distances, the unweighted
start direction, and scripted draws through the rejection loops."""
import math

import numpy as np
import pytest

from akmc_amd import regions as R
from akmc_amd import workloads as W
from oracle import regions_oracle as RO


def _draws(seq):
    it = iter(seq)
    return lambda: next(it)


def test_coordination_closed_forms():
    a = 3.61
    s = W.fcc_sites(a, 4, 4, 4)
    lat = np.diag([4 * a] * 3).reshape(9)
    t = np.zeros(s.shape[1], np.int32)
    nn = a / math.sqrt(2) + 0.1
    assert RO.coordination(s, t, lat, 0, 0, nn) == [12] * s.shape[1]
    c = W.sc_cell(4, 4, 4, n_vac=1, jitter=0.0, seed=2)
    cn = RO.coordination(c.pos, c.types, c.lat, 0, 0, nn)
    d = W.min_image(c.pos - c.vacancies[0][:, None], c.lat)
    shell = np.sqrt((d * d).sum(0)) < nn
    assert shell.sum() == 12 and all(cn[i] == 11 for i in np.nonzero(shell)[0])
    assert all(cn[i] == 12 for i in np.nonzero(~shell)[0])
    assert list(R.coordination_numbers(c.pos, c.types, c.lat, 0, 0, nn)) == cn


def test_rejection_loops_with_scripted_draws():
    n = 10
    frozen = np.zeros(n, np.int32)
    frozen[[3, 4]] = 1
    types = np.array([0, 1, 0, 1, 1, 0, 1, 0, 0, 1], np.int32)
    pos = np.zeros((3, n))
    lat = np.diag([10.0] * 3).reshape(9)
    # style 2: 0.35 -> atom 3 (frozen), 0.41 -> 4 (frozen), 0.55 -> 5
    assert R.choose_centre(2, pos, types, frozen, lat, _draws([0.35, 0.41, 0.55])) == 5
    assert RO.pick([0.35, 0.41, 0.55], n, lambda j: frozen[j] == 0) == 5
    # style 4: one draw for the list entry (0.6 of ["A", "B"] -> "B" = type 1), then atoms: 0.0 -> 0 (type 0),
    # 0.45 -> 4 (frozen), 0.61 -> 6
    j = R.choose_centre(4, pos, types, frozen, lat, _draws([0.6, 0.0, 0.45, 0.61]), species=["A", "B"],
                        art_types=["A", "B"])
    assert j == 6


def test_style5_picks_an_undercoordinated_atom():
    c = W.sc_cell(4, 4, 4, n_vac=1, jitter=0.01, seed=2)
    rng = np.random.default_rng(0)
    nn = c.lat[0] / 4 / math.sqrt(2) + 0.1
    for _ in range(5):
        j = R.choose_centre(5, c.pos, c.types, np.zeros(c.n, np.int32), c.lat, rng.random, species=["Cu"],
                            art_types=["Cu"], coord_distance=[nn], ideal_coord=[12])
        d = W.min_image(c.pos[:, j] - c.vacancies[0], c.lat)
        assert np.sqrt((d * d).sum()) < nn


def test_sphere_and_random_displacement_match_c_oracle(oracle):
    c = W.plain_sc_workload(seed=1)
    rng = np.random.default_rng(3)
    frozen0 = (rng.random(c.n) < 0.1).astype(np.int32)
    j = int(np.nonzero(frozen0 == 0)[0][7])
    fr = R.sphere_region(c.pos, c.lat, j, 6.0, frozen0)
    fo, act = oracle.active_sphere(c.pos, c.lat, j, 6.0, frozen0)
    assert np.array_equal(fr, fo) and act == int((fr == 0).sum())
    nfree = int((fr == 0).sum())
    u = rng.random(3 * nfree)
    disp, mode = R.start_displacement(c.pos, fr, _draws(list(u)), 0.1)
    xo, rv = oracle.random_displacement(c.pos, fr, u, 0.1)
    assert np.allclose(disp[:, fr == 0], (xo - c.pos)[:, fr == 0], atol=1e-15, rtol=0)
    assert np.allclose(mode[:, fr == 0].T.reshape(-1), rv, atol=1e-15, rtol=0)
    assert np.all(disp[:, fr == 1] == 0)


def test_gauss_displacement_weights_and_direction():
    c = W.plain_sc_workload(seed=1)
    rng = np.random.default_rng(4)
    frozen = np.zeros(c.n, np.int32)
    j = 17
    centre = c.pos[:, j].copy()
    u = rng.random(3 * c.n)
    width = 2.5
    disp, mode = R.start_displacement(c.pos, frozen, _draws(list(u)), 0.1, c.lat, centre, width)
    do, ro = RO.gauss_displacement(c.pos, frozen, centre, u, 0.1, width, c.lat)
    assert np.allclose(disp.T, np.array(do), atol=1e-15, rtol=0)
    assert np.allclose(mode.T, np.array(ro), atol=1e-15, rtol=0)
    # closed forms: the centre atom has weight 1; an atom at distance d has exp(-d^2 / 2 w^2);
    # the start direction ignores the weights (it is the plain start's direction)
    plain, pmode = R.start_displacement(c.pos, frozen, _draws(list(u)), 0.1)
    assert np.allclose(disp[:, j], plain[:, j], atol=1e-15)
    k = 40
    d = W.min_image(c.pos[:, k] - centre, c.lat)
    assert np.allclose(disp[:, k], plain[:, k] * math.exp(-(d @ d) / (2 * width ** 2)), atol=1e-15)
    assert np.array_equal(mode, pmode)
