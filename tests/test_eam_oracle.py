"""The tabulated-EAM oracle (oracle/eam_oracle.py), pinned on the CPU:
the setfl reader on hand-written files; exact interpolation of linear
tables (energy equal to its closed form, forces equal to -grad E by
central differences); zero forces on a perfect lattice; and convergence to
the independent Sutton-Chen C oracle when the tables sample the
Sutton-Chen functions.  The library's potentials parser accepts the same
files and rejects what the reference rejects (no device calls here)."""
import ctypes as C
import math

import numpy as np
import pytest

from akmc_amd import workloads as W
from oracle import eam_oracle as E
import eam_fixtures as FX

CU = W.SC_CU


def _linear_tables(tmp_path, dr=0.01, drho=0.01, cut=5.0):
    nr, nrho = int(cut / dr) + 3, 10001
    r = np.arange(nr) * dr
    F = 0.3 - 0.7 * np.arange(nrho) * drho              # F(rho) = 0.3 - 0.7 rho
    rho = np.maximum(2.0 - 0.35 * r, 0.0)               # rho(r) = 2 - 0.35 r (positive inside the cut)
    rphi = r * (1.5 - 0.4 * r)                           # phi(r) = 1.5 - 0.4 r
    p = tmp_path / "lin.eam"
    E.write_lammps(p, nrho, drho, nr, dr, cut, F, rho, rphi, per_line=7)
    return p


def test_reader_layout_and_line_rule(tmp_path):
    text = "\n".join(["c1", "c2", "c3", "1 Cu", "3 0.5 4 0.25 0.7", "29 63.5 3.6 fcc",
                      "1 2", "3 9 9",          # F: 1 2 3, the 9s on F's last line are dropped
                      "4 5 6 7",               # rho
                      "0 0.5 1.0", "1.5"])     # r*phi: 0, 0.5, 1.0, 1.5 -> phi = nan, 2, 2, 2
    t = E.read_lammps(text, is_text=True)
    assert (t.nrho, t.drho, t.nr, t.dr, t.cut) == (3, 0.5, 4, 0.25, 0.7)
    assert list(t.F) == [1, 2, 3] and list(t.rho) == [4, 5, 6, 7]
    assert math.isnan(t.phi[0]) and list(t.phi[1:]) == [2.0, 2.0, 2.0]
    assert list(t.dF) == [2.0, 2.0, 0.0] and list(t.drho_t) == [4.0, 4.0, 4.0, 0.0]


def test_interpolation_operation_order():
    t = np.array([1.0, 3.0, 2.0])
    assert E.interp(t, 0.25, 0.5) == 1.0 + ((3.0 - 1.0) * (0.25 - 0.0)) / 0.5
    assert E.interp(t, 0.75, 0.5) == 3.0 + ((2.0 - 3.0) * (0.75 - 0.5)) / 0.5


def _cell(n=3, a=3.0, jitter=0.05, seed=0):
    rng = np.random.default_rng(seed)
    s = W.fcc_sites(a, n, n, n)
    return s + rng.normal(0, jitter, s.shape), np.zeros(s.shape[1], np.int32), np.diag([a * n] * 3).reshape(9)


def test_linear_tables_closed_form_and_gradient(tmp_path):
    f = E.Field(["Cu"], [("Cu", "Cu", E.read_lammps(_linear_tables(tmp_path)))], rc=4.4)
    pos, types, lat = _cell()
    r = E.forces(f, pos, types, lat)
    # closed form: rho_i = sum_j (2 - 0.35 r), E = sum_i (0.3 - 0.7 rho_i) + 1/2 sum (1.5 - 0.4 r)
    L = np.array([lat[0], lat[4], lat[8]])[:, None]
    n = pos.shape[1]
    e_many, e_pair = 0.0, 0.0
    for a in range(n):
        d = pos[:, a:a + 1] - pos
        d -= L * np.rint(d / L)
        rr = np.sqrt((d * d).sum(0))
        m = (np.arange(n) != a) & (rr <= 4.4)
        e_many += 0.3 - 0.7 * (2.0 - 0.35 * rr[m]).sum()
        e_pair += 0.5 * (1.5 - 0.4 * rr[m]).sum()
    assert r["energy"] == pytest.approx(e_many + e_pair, rel=1e-12)
    # piecewise-linear = linear tables: the forces are exactly -grad E
    h = 1e-5
    for a, c in [(0, 0), (7, 1), (20, 2)]:
        p1, p2 = pos.copy(), pos.copy()
        p1[c, a] += h
        p2[c, a] -= h
        g = (E.forces(f, p1, types, lat)["energy"] - E.forces(f, p2, types, lat)["energy"]) / (2 * h)
        assert r["force"][c, a] == pytest.approx(-g, abs=1e-7)
    assert E.energy_path(f, pos, types, lat) == pytest.approx(r["many"] + 2 * r["pair"], rel=1e-15)


def test_perfect_lattice_has_zero_force(tmp_path):
    # rc below L/2: at exactly L/2 the minimum image keeps one of the two
    # equidistant images, which breaks the symmetry for any potential
    rc = 2 * CU["a"] - 0.05
    p = FX.sc_setfl(tmp_path / "cu.eam", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"], rc, dr=2e-3)
    f = E.Field(["Cu"], [("Cu", "Cu", E.read_lammps(p))], rc=rc)
    s = W.fcc_sites(CU["a"], 4, 4, 4)
    r = E.forces(f, s, np.zeros(s.shape[1], np.int32), np.diag([4 * CU["a"]] * 3).reshape(9))
    assert np.abs(r["force"]).max() < 1e-10


def test_converges_to_sutton_chen_oracle(tmp_path, oracle):
    """Tables sampling the Sutton-Chen functions: energy and forces approach
    the C Sutton-Chen oracle's as the mesh is refined (values O(h^2), the
    forward-difference derivative tables O(h))."""
    c = W.plain_sc_workload(seed=1)
    of = oracle.sc_field(["Cu"], [("Cu", "Cu", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"])], 2 * CU["a"])
    ref = oracle.sc_forces(of, c.pos, c.types, c.lat)
    errs = []
    for h in (4e-3, 1e-3):
        p = FX.sc_setfl(tmp_path / f"cu{h}.eam", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"], 2 * CU["a"],
                        dr=h, drho=h)
        f = E.Field(["Cu"], [("Cu", "Cu", E.read_lammps(p))], rc=2 * CU["a"])
        r = E.forces(f, c.pos, c.types, c.lat)
        errs.append((abs(r["energy"] - ref["energy"]) / c.n, np.abs(r["force"] - ref["force"]).max()))
    print(errs)
    assert errs[1][0] < 1e-6 and errs[1][1] < 2e-3
    assert errs[1][0] < errs[0][0] / 4 and errs[1][1] < errs[0][1] / 2


def test_library_parser_reads_eam_files(tmp_path):
    from akmc_amd import lib
    L = lib.load()
    p = FX.sc_setfl(tmp_path / "cu.eam", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"], 7.22, dr=5e-3, drho=5e-3)
    h = C.c_void_p()
    assert L.akmc_field_create(lib.eam_potentials_text([("Cu", "Cu", str(p))], 7.22).encode(), C.byref(h)) == 0
    assert L.akmc_field_kind(h) == 0
    L.akmc_field_destroy(h)
    bad = [lib.eam_potentials_text([("Cu", "Cu", str(tmp_path / "missing.eam"))], 7.22),
           lib.eam_potentials_text([("Cu", "Cu", str(p))], 7.22).replace("lammps", "dlpoly"),
           lib.eam_potentials_text([("Cu", "Cu", str(p))], 7.22).replace("lammps", "foo")]
    codes = []
    for text in bad:
        codes.append(L.akmc_field_create(text.encode(), C.byref(h)))
    assert codes == [lib.EINVAL, lib.EUNSUPPORTED, lib.EINVAL]
    mixed = lib.eam_potentials_text([("Cu", "Cu", str(p))], 7.22).replace(
        "manybody 1 ev", "manybody 2 ev\nsuttonchen\nCu Cu 0.012382 3.61 9.0 6.0 39.432")
    assert L.akmc_field_create(mixed.encode(), C.byref(h)) == lib.EINVAL
    assert b"can not be mixed" in L.akmc_last_error()
