"""GPU parity of the tabulated EAM field (LAMMPS setfl tables read as
readLEAM does) against the CPU oracle oracle/eam_oracle.py on the same
inputs.  The device interpolates the tables in FP64 on both precision
paths, so the bar is round-off: forces within 1e-9 eV/A, energies and
stress within 1e-11 relative.  Also: two species (three tables, the
per-pair lookup path), the explicit-image path, a device FIRE relax on the
tabulated field, the out-of-table density error, and the energy path's
doubled pair term."""
import numpy as np
import pytest

from akmc_amd import workloads as W
from oracle import eam_oracle as E
import eam_fixtures as FX

pytestmark = pytest.mark.gpu
CU = W.SC_CU


@pytest.fixture(scope="module")
def lib():
    from akmc_amd import lib as L
    L.load()
    return L


def _compare(tag, ref, got, n):
    df = np.abs(got["force"] - ref["force"]).max()
    de = abs(got["energy"]["total"] - ref["energy"])
    ds = np.abs(got["stress"] - ref["stress"]).max() / max(1.0, np.abs(ref["stress"]).max())
    print(f"{tag}: max|dF| {df:.2e} eV/A, |dE| {de:.2e} eV, stress rel {ds:.2e}", flush=True)
    assert df < 1e-9
    assert de < 1e-11 * abs(ref["energy"])
    assert abs(got["energy"]["many"] - ref["many"]) < 1e-11 * abs(ref["many"])
    assert ds < 1e-11
    assert np.abs(got["pos"] - ref["pos"]).max() < 1e-12


@pytest.mark.parametrize("bits", [32, 64])
def test_eam_forces_match_oracle(lib, tmp_path, bits):
    c = W.plain_sc_workload(seed=1)
    rc = 2 * CU["a"]
    p = FX.sc_setfl(tmp_path / "cu.eam", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"], rc)
    f = lib.Field(lib.eam_potentials_text([("Cu", "Cu", str(p))], rc))
    f.set_precision(bits)
    of = E.Field(["Cu"], [("Cu", "Cu", E.read_lammps(p))], rc)
    ref = E.forces(of, c.pos, c.types, c.lat)
    got = f.calculate_forces(c.pos, c.types, c.lat)
    _compare(f"EAM C1 fp{bits}", ref, got, c.n)
    e = f.calculate_energy(c.pos, c.types, c.lat)
    assert abs(e["total"] - E.energy_path(of, c.pos, c.types, c.lat)) < 1e-11 * abs(ref["energy"])


def test_eam_two_species_tables(lib, tmp_path):
    c = W.sc_cell(5, 5, 5, n_vac=2, jitter=0.05, seed=7, two_species=True, a=3.7)
    rc = 7.3
    pots = {("A", "A"): (0.0124, 3.61, 9.0, 6.0, 39.4), ("B", "B"): (0.0158, 3.89, 10.0, 8.0, 34.4),
            ("A", "B"): (0.0140, 3.75, 9.5, 7.0, 36.8)}
    entries = []
    for (A, B), (eps, a, n, m, cc) in pots.items():
        p = FX.sc_setfl(tmp_path / f"{A}{B}.eam", eps, a, n, m, cc, rc, dr=2e-3, drho=1e-2, rho_max=300.0)
        entries.append((A, B, str(p)))
    f = lib.Field(lib.eam_potentials_text(entries, rc, species=["A", "B"]))
    of = E.Field(["A", "B"], [(A, B, E.read_lammps(p)) for A, B, p in entries], rc)
    ref = E.forces(of, c.pos, c.types, c.lat)
    got = f.calculate_forces(c.pos, c.types, c.lat)
    _compare(f"EAM two species {c.n} atoms", ref, got, c.n)


def test_eam_explicit_images(lib, tmp_path):
    c = W.sc_cell(3, 3, 3, n_vac=1, jitter=0.02, seed=3)
    rc = 2 * CU["a"]
    p = FX.sc_setfl(tmp_path / "cu.eam", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"], rc, dr=2e-3, drho=2e-3)
    f = lib.Field(lib.eam_potentials_text([("Cu", "Cu", str(p))], rc, noimage=True))
    of = E.Field(["Cu"], [("Cu", "Cu", E.read_lammps(p))], rc, min_image=False)
    ref = E.forces(of, c.pos, c.types, c.lat)
    got = f.calculate_forces(c.pos, c.types, c.lat)
    _compare(f"EAM no-image {c.n} atoms", ref, got, c.n)


@pytest.mark.parametrize("case", ["one_table", "two_species", "noimage"])
def test_eam_lds_staged(lib, tmp_path, case):
    """The LDS-staged EAM kernels (a walker's gather records in LDS, the
    tables in global memory) give the unstaged kernels' forces and energies
    bitwise (same FP64 operations in the same order) and the oracle's to
    round-off, on the one-table, per-pair-table and explicit-image paths."""
    if case == "two_species":
        cells = [W.sc_cell(5, 5, 5, n_vac=2, jitter=0.05, seed=7 + k, two_species=True, a=3.7) for k in range(2)]
        rc = 7.3
        pots = {("A", "A"): (0.0124, 3.61, 9.0, 6.0, 39.4), ("B", "B"): (0.0158, 3.89, 10.0, 8.0, 34.4),
                ("A", "B"): (0.0140, 3.75, 9.5, 7.0, 36.8)}
        entries = []
        for (A, B), (eps, a, n, m, cc) in pots.items():
            p = FX.sc_setfl(tmp_path / f"{A}{B}.eam", eps, a, n, m, cc, rc, dr=2e-3, drho=1e-2, rho_max=300.0)
            entries.append((A, B, str(p)))
        text = lib.eam_potentials_text(entries, rc, species=["A", "B"])
        of = E.Field(["A", "B"], [(A, B, E.read_lammps(p)) for A, B, p in entries], rc)
    else:
        rc = 2 * CU["a"]
        p = FX.sc_setfl(tmp_path / "cu.eam", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"], rc, dr=2e-3, drho=2e-3)
        if case == "noimage":
            cells = [W.sc_cell(3, 3, 3, n_vac=1, jitter=0.02, seed=3 + k) for k in range(3)]
            text = lib.eam_potentials_text([("Cu", "Cu", str(p))], rc, noimage=True)
        else:
            cells = [W.plain_sc_workload(seed=1), W.sc_cell(5, 5, 5, n_vac=2, jitter=0.03, seed=9)]
            text = lib.eam_potentials_text([("Cu", "Cu", str(p))], rc)
        of = E.Field(["Cu"], [("Cu", "Cu", E.read_lammps(p))], rc, min_image=case != "noimage")
    f = lib.Field(text)
    b = lib.Batch(f, cells)
    assert b.staging() == (False, True)
    e0 = b.forces()
    _, f0 = b.download()
    b.set_staging(True)
    e1 = b.forces()
    _, f1 = b.download()
    for w, c in enumerate(cells):
        ref = E.forces(of, c.pos, c.types, c.lat)
        df = np.abs(f1[w] - ref["force"]).max()
        print(f"staged EAM {case} walker {w} ({c.n} atoms): max|dF| {df:.2e} eV/A", flush=True)
        assert df < 1e-9 and abs(e1[w]["total"] - ref["energy"]) < 1e-11 * abs(ref["energy"])
        assert np.array_equal(f1[w], f0[w]) and e1[w]["total"] == e0[w]["total"]


def test_eam_relax_and_dimer_on_device(lib, tmp_path):
    """FIRE relax and a dimer search on the tabulated field: the oracle's
    forces at the device minimum are below the relax tolerance, and the
    saddle's FP64 barrier matches the Sutton-Chen hop it samples
    (0.536 eV) to the tables' accuracy."""
    c = W.plain_sc_workload(seed=1)
    rc = 2 * CU["a"]
    p = FX.sc_setfl(tmp_path / "cu.eam", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"], rc)
    f = lib.Field(lib.eam_potentials_text([("Cu", "Cu", str(p))], rc))
    of = E.Field(["Cu"], [("Cu", "Cu", E.read_lammps(p))], rc)
    b = lib.Batch(f, [c])
    b.start_relax()
    b.run(5000, sync_every=25)
    st = b.status()[0]
    basin = b.positions()[0]
    r1 = E.forces(of, basin, c.types, c.lat)
    fmax = np.sqrt((r1["force"] ** 2).sum(0).max())
    print(f"EAM relax: status {st['status']} iters {st['iterations']} E {st['energy']:.8f} "
          f"(oracle {r1['energy']:.8f}) oracle |F|max {fmax:.2e}", flush=True)
    assert st["status"] == 0 and fmax < 2e-4
    assert abs(st["energy"] - r1["energy"]) < 1e-9 * abs(r1["energy"])
    dv = W.min_image(basin - c.vacancies[0][:, None], c.lat)
    frozen = (np.sqrt((dv * dv).sum(0)) >= 6.0).astype(np.int32)
    nb = W.neighbours_of_site(W.Cell(basin, c.types, c.lat), c.vacancies[0], 12)
    disp, mode = W.local_displacement(W.Cell(basin, c.types, c.lat), nb, 0.1, 3)
    b.set_verify(16, 0.01)
    b.start_dimer([mode], params=lib.dimer_defaults(), x0_per_walker=[basin + disp], frozen_per_walker=[frozen])
    b.run(40000, sync_every=25)
    s = b.status()[0]
    sad = b.positions()[0]
    e2 = E.forces(of, sad, c.types, c.lat)["energy"]
    print(f"EAM dimer: status {s['status']} verified {s['verified']} Ea {e2 - r1['energy']:.5f} "
          f"lambdas {s['lambda1']:.4f} {s['lambda2']:.4f}", flush=True)
    assert s["status"] == 0 and s["verified"] == 1 and s["lambda1"] < 0 < s["lambda2"]
    assert abs(s["energy64"] - e2) < 1e-9 * abs(e2)
    assert abs((e2 - r1["energy"]) - 0.536) < 0.01


def test_eam_density_outside_table_is_an_error(lib, tmp_path):
    c = W.plain_sc_workload(seed=1)
    rc = 2 * CU["a"]
    p = FX.sc_setfl(tmp_path / "short.eam", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"], rc, rho_max=10.0)
    f = lib.Field(lib.eam_potentials_text([("Cu", "Cu", str(p))], rc))
    with pytest.raises(lib.AkmcError, match="out of bounds"):
        f.calculate_forces(c.pos, c.types, c.lat)
