"""Pins of the CPU FP64 oracle's force fields (no GPU).

What pins what:
  * Sutton-Chen (potentials/metal/): perfect-FCC forces vanish by symmetry;
    perfect-lattice energy equals the textbook Sutton-Chen lattice sum
    eps[1/2 S_n - c sqrt(S_m)] (Sutton & Chen 1990) computed here with numpy;
    an independent numpy statement of E = eps sum_i [1/2 sum_j (a/r)^n -
    c sqrt(sum_j (a/r)^m)] on a two-species cell; F = -grad E by central
    differences; sum F = 0; translation invariance; image and no-image paths
    agree when rc < L/2; the cutoff is r <= rc (an atom pair exactly at rc
    counts).
  * Ewald + Buckingham (potentials/pair/): the reference-held energy of
    docs/examples/gdc (2549 ions, -87780.5106508483 eV, reproduced as a
    fixture under tests/golden); the rock-salt Madelung constant 1.747565;
    eta-independence of the total; the k-set size for |n| < 8; a single
    Buckingham pair against its closed form; F = -grad E; sum F = 0.
  * tridiagonal eigensolver: numpy.linalg.eigh on random tridiagonals.
"""
import json
import os

import numpy as np
import pytest

from akmc_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CU = W.SC_CU


def _sc(O, rc=2 * CU["a"], min_image=True):
    return O.sc_field(["Cu"], [("Cu", "Cu", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"])], rc=rc,
                      min_image=min_image)


def test_sc_perfect_lattice_closed_form(oracle):
    f = _sc(oracle, rc=2 * CU["a"] - 1e-6)
    c = W.sc_cell(4, 4, 4, n_vac=0, jitter=0)
    r = oracle.sc_forces(f, c.pos, c.types, c.lat)
    assert np.abs(r["force"]).max() < 1e-12
    # textbook lattice sums over a large FCC block
    s = W.fcc_sites(CU["a"], 10, 10, 10)
    d = W.min_image(s - s[:, :1], np.diag([10 * CU["a"]] * 3).reshape(9))
    rr = np.sqrt((d * d).sum(0))
    rr = rr[(rr > 0) & (rr <= 2 * CU["a"] - 1e-6)]
    Sn = ((CU["a"] / rr) ** CU["n"]).sum()
    Sm = ((CU["a"] / rr) ** CU["m"]).sum()
    e_atom = CU["eps"] * (0.5 * Sn - CU["c"] * np.sqrt(Sm))
    assert abs(r["energy"] / c.n - e_atom) < 1e-12 * abs(e_atom)
    assert abs(Sm - 113.5629409351) < 1e-8  # 134 neighbours inside 2a


def test_sc_cutoff_is_inclusive(oracle):
    """rc = 2a exactly: the 2a shell (neighbours at r = rc) must count,
    Field.cpp:587 tests rsq <= radius."""
    c = W.sc_cell(4, 4, 4, n_vac=0, jitter=0)
    f_in = _sc(oracle, rc=2 * CU["a"])
    f_out = _sc(oracle, rc=2 * CU["a"] - 1e-9)
    n_in = oracle.sc_count_pairs(f_in, c.pos, c.lat)
    n_out = oracle.sc_count_pairs(f_out, c.pos, c.lat)
    assert n_out == 256 * 134
    # with L = 4a the six (2a,0,0)-type sites are three minimum images at
    # exactly r = rc = L/2, all counted by the inclusive test: 137 each
    assert n_in == 256 * 137


def _numpy_sc_energy(pos, types, lat, pars, rc):
    """Sutton-Chen energy from its textbook form with per-pair parameters
    pars[(ti,tj)] = (eps, a, n, m, c)."""
    n = pos.shape[1]
    d = pos[:, :, None] - pos[:, None, :]
    d = W.min_image(d, lat)
    r = np.sqrt((d * d).sum(0))
    e = 0.0
    for i in range(n):
        rep = 0.0
        dens = 0.0
        for j in range(n):
            if i == j or r[i, j] > rc:
                continue
            eps, a, nn, m, c = pars[(types[i], types[j])]
            rep += eps * (a / r[i, j]) ** nn
            dens += (eps * c) ** 2 * (a / r[i, j]) ** m
        e += 0.5 * rep - np.sqrt(dens)
    return e


def test_sc_two_species_textbook_energy(oracle):
    c = W.sc_cell(3, 3, 3, n_vac=2, jitter=0.05, seed=7, two_species=True, a=3.7)
    pa = (0.0124, 3.61, 9.0, 6.0, 39.4)
    pb = (0.0158, 3.89, 10.0, 8.0, 34.4)
    pab = (0.0140, 3.75, 9.5, 7.0, 36.8)
    f = oracle.sc_field(["A", "B"], [("A", "A") + pa, ("B", "B") + pb, ("A", "B") + pab], rc=5.5)
    r = oracle.sc_forces(f, c.pos, c.types, c.lat)
    pars = {(0, 0): pa, (1, 1): pb, (0, 1): pab, (1, 0): pab}
    e_ref = _numpy_sc_energy(r["pos"], c.types, c.lat, pars, 5.5)
    assert abs(r["energy"] - e_ref) < 1e-11 * abs(e_ref)


def _fd_check(energy_fn, pos, force, atoms, h=1e-5, tol=2e-6):
    for i in atoms:
        for k in range(3):
            p = pos.copy()
            p[k, i] += h
            ep = energy_fn(p)
            p[k, i] -= 2 * h
            em = energy_fn(p)
            fd = -(ep - em) / (2 * h)
            assert abs(fd - force[k, i]) < tol, (i, k, fd, force[k, i])


def test_sc_forces_are_gradient(oracle):
    f = _sc(oracle)
    c = W.plain_sc_workload(seed=1)
    r = oracle.sc_forces(f, c.pos, c.types, c.lat)
    assert np.abs(r["force"].sum(1)).max() < 1e-11
    _fd_check(lambda p: oracle.sc_energy(f, p, c.types, c.lat)[0], r["pos"], r["force"], [0, 17, 100, 254])


def test_sc_translation_and_image_paths(oracle):
    c = W.sc_cell(5, 5, 5, n_vac=1, jitter=0.03, seed=4)
    f = _sc(oracle, rc=2 * CU["a"])
    fni = _sc(oracle, rc=2 * CU["a"], min_image=False)
    a = oracle.sc_forces(f, c.pos, c.types, c.lat)
    b = oracle.sc_forces(f, c.pos + np.array([[0.37], [-1.2], [5.1]]), c.types, c.lat)
    d = oracle.sc_forces(fni, c.pos, c.types, c.lat)
    assert abs(a["energy"] - b["energy"]) < 1e-9
    assert np.abs(a["force"] - b["force"]).max() < 1e-11
    assert abs(a["energy"] - d["energy"]) < 1e-9
    assert np.abs(a["force"] - d["force"]).max() < 1e-11


def _gdc(oracle):
    with open(os.path.join(GOLDEN, "gdc_example.json")) as fh:
        g = json.load(fh)
    return g


def test_ewald_gdc_reference_energy(oracle):
    """docs/examples/gdc/basis:3 holds -87780.5106508483 eV for the cell in
    that file with the potentials of docs/examples/gdc/potentials."""
    g = _gdc(oracle)
    lat = np.array(g["lat"])
    pos = np.array(g["pos"])
    ew = oracle.Ewald(lat, None, acc=g["ewald_precis"], rc=g["cutoff"])
    assert ew.gcell == [8, 8, 8] and ew.nk == 1147
    pots = [tuple(p) for p in g["buckingham"]]
    f = oracle.ion_field(g["species"], pots, rc=g["cutoff"], ewald=ew, min_image=True, lat=lat)
    r = oracle.ion_forces(f, pos, np.array(g["types"]), np.array(g["charges"]), lat)
    assert abs(r["energy"] - g["energy"]) < 1e-6, r["energy"] - g["energy"]


def test_ewald_madelung_constant(oracle):
    c = W.rocksalt_cell(jitter=0, schottky=False, charge=1.0)
    ew = oracle.Ewald(c.lat, 0.4, gcell=10)
    f = oracle.ion_field(["M", "X"], [], rc=10.0, ewald=ew, min_image=False, lat=c.lat, table_default=-1)
    r = oracle.ion_forces(f, c.pos, c.types, c.charges, c.lat)
    M = -r["energy"] / (c.n / 2) / oracle.CTOEV * (4.212 / 2)
    assert abs(M - 1.747564594633) < 2e-7
    assert np.abs(r["force"]).max() < 1e-10


def test_ewald_kvector_count(oracle):
    lat = np.diag([16.848] * 3).reshape(9)
    ew = oracle.Ewald(lat, 0.35, gcell=8)
    n = np.arange(-8, 9)
    g = np.stack(np.meshgrid(n, n, n, indexing="ij"), -1).reshape(-1, 3)
    s = (g * g).sum(1)
    half = ((g[:, 0] > 0) | (g[:, 0] == 0)) & (s > 0) & (s < 64)
    assert ew.nk == int(half.sum()) == 1147


def _c2(oracle, eta=0.35, g=8, seed=5):
    c = W.rocksalt_cell(seed=seed)
    ew = oracle.Ewald(c.lat, eta, gcell=g)
    pots = [("M", "X") + W.ROCKSALT_BUCK[("M", "X")], ("X", "X") + W.ROCKSALT_BUCK[("X", "X")],
            ("M", "M") + W.ROCKSALT_BUCK[("M", "M")]]
    f = oracle.ion_field(["M", "X"], pots, rc=10.0, ewald=ew, min_image=False, lat=c.lat)
    return c, f


def test_ewald_eta_independence(oracle):
    es = []
    # converged (eta, |n|max) pairs at rc = 10; the workload's own eta = 0.35
    # carries ~1e-2 eV of real-space truncation at this cutoff
    for eta, g in [(0.4, 10), (0.45, 12), (0.5, 14)]:
        c, f = _c2(oracle, eta, g)
        es.append(oracle.ion_forces(f, c.pos, c.types, c.charges, c.lat)["energy"])
    assert max(es) - min(es) < 5e-4


@pytest.mark.slow
def test_ion_forces_are_gradient(oracle):
    c, f = _c2(oracle)
    r = oracle.ion_forces(f, c.pos, c.types, c.charges, c.lat)
    assert np.abs(r["force"].sum(1)).max() < 1e-9
    _fd_check(lambda p: oracle.ion_forces(f, p, c.types, c.charges, c.lat)["energy"], r["pos"], r["force"],
              [3, 300], h=1e-5, tol=5e-6)


def test_buckingham_single_pair(oracle):
    lat = np.diag([40.0] * 3).reshape(9)
    pos = np.array([[0.0, 2.3], [0.0, 0.4], [0.0, -0.1]])
    A, rho, Cc = 22764.0, 0.149, 27.88
    f = oracle.ion_field(["X"], [("X", "X", A, rho, Cc)], rc=10.0, ewald=None, min_image=True)
    r = oracle.ion_forces(f, pos, np.array([0, 0]), np.zeros(2), lat)
    rr = np.sqrt(2.3 ** 2 + 0.4 ** 2 + 0.1 ** 2)
    assert abs(r["energy"] - (A * np.exp(-rr / rho) - Cc / rr ** 6)) < 1e-12
    dVdr = -A / rho * np.exp(-rr / rho) + 6 * Cc / rr ** 7
    u = (pos[:, 0] - pos[:, 1]) / rr
    assert np.abs(r["force"][:, 0] - (-dVdr * u)).max() < 1e-12


def test_tridiag_eigensolver(oracle):
    rng = np.random.default_rng(11)
    for n in [1, 2, 3, 7, 20, 25]:
        d = rng.normal(size=n)
        e = rng.normal(size=n - 1)
        w, z = oracle.tridiag_eig(d, e)
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        wr, zr = np.linalg.eigh(T)
        assert np.abs(w - wr).max() < 1e-12
        for k in range(n):
            assert abs(abs(z[:, k] @ zr[:, k]) - 1.0) < 1e-10
        assert np.abs(T @ z - z * w).max() < 1e-12
