"""GPU parity of the force fields: HIP kernels through the C ABI vs the FP64
oracle on identical seeded inputs.

Tolerances (the north-star budget): forces within 1e-5 eV/A of FP64,
energies within 1e-6 eV/atom; the FP64 verification path within 1e-9.
The measured error against FP64 is printed for every case.
"""
import numpy as np
import pytest

from akmc_amd import workloads as W

pytestmark = pytest.mark.gpu

CU = W.SC_CU
F_TOL = 1e-5
E_TOL_ATOM = 1e-6


def _sc_pots():
    return [("Cu", "Cu", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"])]


@pytest.fixture(scope="module")
def lib():
    from akmc_amd import lib as L
    L.load()
    return L


def _compare(name, ref, got, n, ftol=F_TOL, etol=E_TOL_ATOM):
    df = np.abs(got["force"] - ref["force"]).max()
    de = abs(got["energy"]["total"] - ref["energy"]) / n
    print(f"{name}: max|dF| = {df:.3e} eV/A, |dE|/N = {de:.3e} eV/atom", flush=True)
    assert df < ftol, df
    assert de < etol, de
    assert np.abs(got["pos"] - ref["pos"]).max() < 1e-9  # wrapped like the reference


@pytest.mark.parametrize("fp64", [False, True])
def test_sc_plain_cell(lib, oracle, fp64):
    c = W.plain_sc_workload(seed=1)
    f = lib.Field(lib.sc_potentials_text(_sc_pots(), 2 * CU["a"]))
    if fp64:
        f.set_precision(64)
    ref = oracle.sc_forces(oracle.sc_field(["Cu"], _sc_pots(), 2 * CU["a"]), c.pos, c.types, c.lat)
    got = f.calculate_forces(c.pos, c.types, c.lat)
    tol = (1e-9, 1e-11) if fp64 else (F_TOL, E_TOL_ATOM)
    _compare(f"SC C1 fp{64 if fp64 else 32}", ref, got, c.n, *tol)
    assert abs(got["energy"]["many"] - ref["e_many"]) / c.n < tol[1]
    # stress is a cancelling sum of ~3.5e4 O(1 eV) pair terms: FP32 pair terms
    # leave ~1e-5 relative; the FP64 path must agree to round-off
    stol = 1e-9 if fp64 else 1e-4
    ds = np.abs(got["stress"] - ref["stress"]).max() / max(1.0, np.abs(ref["stress"]).max())
    print(f"  stress rel err {ds:.2e}", flush=True)
    assert ds < stol


def test_sc_two_species_and_large_cell(lib, oracle):
    c = W.sc_cell(9, 9, 9, n_vac=3, jitter=0.05, seed=7, two_species=True, a=3.7)
    pots = [("A", "A", 0.0124, 3.61, 9.0, 6.0, 39.4), ("B", "B", 0.0158, 3.89, 10.0, 8.0, 34.4),
            ("A", "B", 0.0140, 3.75, 9.5, 7.0, 36.8)]
    f = lib.Field(lib.sc_potentials_text(pots, 7.3, species=["A", "B"]))
    ref = oracle.sc_forces(oracle.sc_field(["A", "B"], pots, 7.3), c.pos, c.types, c.lat)
    got = f.calculate_forces(c.pos, c.types, c.lat)
    _compare("SC two-species 2913 atoms", ref, got, c.n)


def test_sc_noimage_path(lib, oracle):
    c = W.sc_cell(3, 3, 3, n_vac=1, jitter=0.02, seed=3)
    f = lib.Field(lib.sc_potentials_text(_sc_pots(), 7.0, noimage=True))
    ref = oracle.sc_forces(oracle.sc_field(["Cu"], _sc_pots(), 7.0, min_image=False), c.pos, c.types, c.lat)
    got = f.calculate_forces(c.pos, c.types, c.lat)
    _compare("SC no-image (rc > L/2)", ref, got, c.n)


def _c2_field(lib, oracle, c, eta=0.35, g=8):
    pots = [("M", "X") + W.ROCKSALT_BUCK[("M", "X")], ("X", "X") + W.ROCKSALT_BUCK[("X", "X")],
            ("M", "M") + W.ROCKSALT_BUCK[("M", "M")]]
    f = lib.Field(lib.buck_potentials_text(["M", "X"], [2.0, -2.0], pots, 10.0, noimage=True))
    f.set_ewald(eta, g)
    ew = oracle.Ewald(c.lat, eta, gcell=g)
    of = oracle.ion_field(["M", "X"], pots, rc=10.0, ewald=ew, min_image=False, lat=c.lat)
    return f, of


@pytest.mark.parametrize("fp64", [False, True])
def test_ionic_rocksalt_ewald(lib, oracle, fp64):
    c = W.rocksalt_cell(seed=5)
    f, of = _c2_field(lib, oracle, c)
    if fp64:
        f.set_precision(64)
    ref = oracle.ion_forces(of, c.pos, c.types, c.charges, c.lat)
    got = f.calculate_forces(c.pos, c.types, c.lat, c.charges)
    assert f.num_kvectors == 1147
    tol = (1e-8, 1e-10) if fp64 else (F_TOL, E_TOL_ATOM)
    _compare(f"rock-salt Ewald fp{64 if fp64 else 32}", ref, got, c.n, *tol)
    for k, rk in (("rcp", "e_rcp"), ("real", "e_real"), ("vdw", "e_vdw")):
        assert abs(got["energy"][k] - ref[rk]) / c.n < tol[1], k


def test_ionic_gdc_reference_energy(lib, oracle):
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "gdc_example.json")))
    pots = [tuple(p) for p in g["buckingham"]]
    sp = g["species"]
    q = {s: None for s in sp}
    for t, qq in zip(g["types"], g["charges"]):
        q[sp[t]] = qq
    f = lib.Field(lib.buck_potentials_text(sp, [q[s] for s in sp], pots, g["cutoff"], ewald_precis=1e-6))
    got = f.calculate_forces(np.array(g["pos"]), np.array(g["types"]), np.array(g["lat"]), np.array(g["charges"]))
    n = len(g["types"])
    de = abs(got["energy"]["total"] - g["energy"]) / n
    print(f"gdc 2549 ions: |E - E_ref|/N = {de:.3e} eV/atom", flush=True)
    assert de < E_TOL_ATOM


def test_batch_matches_single(lib, oracle):
    cells = W.batched_sc_cells(n_walkers=6, seed=11)
    f = lib.Field(lib.sc_potentials_text(_sc_pots(), 2 * CU["a"]))
    b = lib.Batch(f, cells)
    e = b.forces()
    _, frc = b.download()
    of = oracle.sc_field(["Cu"], _sc_pots(), 2 * CU["a"])
    for w, c in enumerate(cells):
        ref = oracle.sc_forces(of, c.pos, c.types, c.lat)
        df = np.abs(frc[w] - ref["force"]).max()
        de = abs(e[w]["total"] - ref["energy"]) / c.n
        print(f"walker {w} N={c.n}: max|dF| {df:.2e} |dE|/N {de:.2e}", flush=True)
        assert df < F_TOL and de < E_TOL_ATOM


def test_skin_reuse_matches_fresh(lib, oracle):
    """Forces after small moves (list reused inside the skin) equal the
    oracle's; a move beyond half the skin triggers a rebuild."""
    c = W.plain_sc_workload(seed=2)
    f = lib.Field(lib.sc_potentials_text(_sc_pots(), 2 * CU["a"]))
    b = lib.Batch(f, [c])
    b.forces()
    r0 = b.rebuilds
    rng = np.random.default_rng(0)
    of = oracle.sc_field(["Cu"], _sc_pots(), 2 * CU["a"])
    pos = c.pos.copy()
    for step in range(3):
        pos = pos + rng.normal(0, 0.05, pos.shape)
        c2 = W.Cell(pos, c.types, c.lat)
        b.upload([c2])
        e = b.forces()
        _, frc = b.download()
        ref = oracle.sc_forces(of, pos, c.types, c.lat)
        assert np.abs(frc[0] - ref["force"]).max() < F_TOL
        assert abs(e[0]["total"] - ref["energy"]) / c.n < E_TOL_ATOM
    assert b.rebuilds >= r0


@pytest.mark.parametrize("fp64", [False, True])
def test_kernel_variants_and_autotune(lib, oracle, fp64):
    """Every (lanes per atom, gathers in flight) variant of the Sutton-Chen
    kernels gives the oracle's forces (FP32 within the force budget, FP64 to
    round-off); variants with the same lanes are bitwise identical; the
    autotuner keeps one of them without changing its results."""
    cells = W.batched_sc_cells(n_walkers=5, seed=13)
    f = lib.Field(lib.sc_potentials_text(_sc_pots(), 2 * CU["a"]))
    if fp64:
        f.set_precision(64)
    b = lib.Batch(f, cells)
    of = oracle.sc_field(["Cu"], _sc_pots(), 2 * CU["a"])
    refs = [oracle.sc_forces(of, c.pos, c.types, c.lat) for c in cells]
    ftol, etol = (1e-9, 1e-11) if fp64 else (F_TOL, E_TOL_ATOM)
    got = {}
    for lanes in (1, 2, 4, 8):
        for unroll in (1, 2, 4, 8):
            b.set_variant(lanes, unroll)
            e = b.forces()
            _, frc = b.download()
            got[(lanes, unroll)] = [f.copy() for f in frc]
            worst = 0.0
            for w, c in enumerate(cells):
                df = np.abs(frc[w] - refs[w]["force"]).max()
                de = abs(e[w]["total"] - refs[w]["energy"]) / c.n
                worst = max(worst, df)
                assert df < ftol and de < etol, (lanes, unroll, w, df, de)
            print(f"variant {lanes}x{unroll} fp{64 if fp64 else 32}: max|dF| {worst:.2e}", flush=True)
        for w in range(len(cells)):
            for unroll in (2, 4, 8):
                # same summation order; the FP32 path is bitwise equal, the FP64
                # path may differ by the compiler's FMA contraction choices
                if fp64:
                    assert np.abs(got[(lanes, 1)][w] - got[(lanes, unroll)][w]).max() < 1e-12
                else:
                    assert np.array_equal(got[(lanes, 1)][w], got[(lanes, unroll)][w])
    # LDS-staged variants: the same per-lane summation order, so bitwise equal
    # to the unstaged ones in FP32 (to FMA contraction in FP64)
    assert b.staging() == (False, True)
    b.set_staging(True)
    for lanes in (1, 2, 4, 8):
        for unroll in (1, 4):
            b.set_variant(lanes, unroll)
            b.forces()
            _, frc = b.download()
            for w in range(len(cells)):
                if fp64:
                    assert np.abs(frc[w] - got[(lanes, unroll)][w]).max() < 1e-12, (lanes, unroll, w)
                else:
                    assert np.array_equal(frc[w], got[(lanes, unroll)][w]), (lanes, unroll, w)
    b.set_staging(False)
    best, ms = b.autotune(nrep=5)
    print(f"autotune: best {best}, ms per evaluation {ms}", flush=True)
    assert best[:2] in got and len(ms) == 32
    assert b.staging()[0] == bool(best[2])
    b.forces()
    _, frc = b.download()
    for w in range(len(cells)):
        if fp64 and best[2]:
            assert np.abs(frc[w] - got[best[:2]][w]).max() < 1e-12
        else:
            assert np.array_equal(frc[w], got[best[:2]][w])
    with pytest.raises(Exception):
        b.set_variant(3, 1)
    with pytest.raises(Exception):
        b.set_variant(1, 3)


@pytest.mark.parametrize("fp64", [False, True])
def test_sc_non_integer_exponents(lib, oracle, fp64):
    """Sutton-Chen with non-integer n and m (the general power path: libm
    pow in FP64, hardware log2/exp2 in FP32) against the oracle."""
    pots = [("Cu", "Cu", CU["eps"], CU["a"], 8.7, 5.6, CU["c"])]
    c = W.plain_sc_workload(seed=4)
    f = lib.Field(lib.sc_potentials_text(pots, 2 * CU["a"]))
    if fp64:
        f.set_precision(64)
    ref = oracle.sc_forces(oracle.sc_field(["Cu"], pots, 2 * CU["a"]), c.pos, c.types, c.lat)
    got = f.calculate_forces(c.pos, c.types, c.lat)
    tol = (1e-9, 1e-11) if fp64 else (F_TOL, E_TOL_ATOM)
    _compare(f"SC n=8.7 m=5.6 fp{64 if fp64 else 32}", ref, got, c.n, *tol)


_TWO_SPECIES = [("A", "A", 0.0124, 3.61, 9.0, 6.0, 39.4), ("B", "B", 0.0158, 3.89, 10.0, 8.0, 34.4),
                ("A", "B", 0.0140, 3.75, 9.5, 7.0, 36.8)]


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("case", ["two_species", "noimage", "non_integer"])
def test_lds_staged_kernels(lib, oracle, case, fp64):
    """The LDS-staged Sutton-Chen kernels (a walker's gather records copied
    to LDS by its workgroups) on the multi-species, explicit-image and
    non-integer-exponent paths: the oracle's forces and energies, and the
    unstaged kernels' results bitwise in FP32 (to FMA contraction in FP64)."""
    if case == "two_species":
        cells = [W.sc_cell(9, 9, 9, n_vac=3, jitter=0.05, seed=7 + k, two_species=True, a=3.7) for k in range(2)]
        text = lib.sc_potentials_text(_TWO_SPECIES, 7.3, species=["A", "B"])
        of = oracle.sc_field(["A", "B"], _TWO_SPECIES, 7.3)
    elif case == "noimage":
        cells = [W.sc_cell(3, 3, 3, n_vac=1, jitter=0.02, seed=3 + k) for k in range(3)]
        text = lib.sc_potentials_text(_sc_pots(), 7.0, noimage=True)
        of = oracle.sc_field(["Cu"], _sc_pots(), 7.0, min_image=False)
    else:
        pots = [("Cu", "Cu", CU["eps"], CU["a"], 8.7, 5.6, CU["c"])]
        cells = W.batched_sc_cells(n_walkers=3, seed=21)
        text = lib.sc_potentials_text(pots, 2 * CU["a"])
        of = oracle.sc_field(["Cu"], pots, 2 * CU["a"])
    f = lib.Field(text)
    if fp64:
        f.set_precision(64)
    b = lib.Batch(f, cells)
    assert b.staging() == (False, True)
    ftol, etol = (1e-9, 1e-11) if fp64 else (F_TOL, E_TOL_ATOM)
    for lanes, unroll in ((1, 1), (4, 4), (8, 2)):
        b.set_variant(lanes, unroll)
        b.set_staging(False)
        e0 = b.forces()
        _, f0 = b.download()
        b.set_staging(True)
        e1 = b.forces()
        _, f1 = b.download()
        for w, c in enumerate(cells):
            ref = oracle.sc_forces(of, c.pos, c.types, c.lat)
            df = np.abs(f1[w] - ref["force"]).max()
            de = abs(e1[w]["total"] - ref["energy"]) / c.n
            print(f"staged {case} {lanes}x{unroll} fp{64 if fp64 else 32} walker {w}: max|dF| {df:.2e}, "
                  f"|dE|/N {de:.2e}", flush=True)
            assert df < ftol and de < etol
            if fp64:
                assert np.abs(f1[w] - f0[w]).max() < 1e-12
            else:
                assert np.array_equal(f1[w], f0[w]) and e1[w]["total"] == e0[w]["total"]


_FORMS = {
    "morse_bhm_lj": [("morse", "M", "X", 0.6, 1.6, 2.1, 4.0), ("bhm", "X", "X", 1200.0, 3.0, 0.9, 20.0, 8.0),
                     ("ljones", "M", "M", 0.01, 2.5)],
    "sevnsix": [("severnsix", "M", "X", 0.3, 2.2), ("buck", "X", "X", 22764.0, 0.149, 27.88),
                ("severnsix", "M", "M", 0.05, 3.0)],
}


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("forms", sorted(_FORMS))
def test_ionic_other_pair_forms(lib, oracle, forms, fp64):
    """Morse, Born-Huggins-Mayer, Lennard-Jones and 7-6 short-range forms
    with Ewald, as TwoBody::shortrangeone computes them, against the
    oracle; the energy-only entry point adds Morse's C12/r^12 term."""
    c = W.rocksalt_cell(seed=5)
    pots = _FORMS[forms]
    f = lib.Field(lib.buck_potentials_text(["M", "X"], [2.0, -2.0], pots, 10.0, noimage=True))
    f.set_ewald(0.35, 8)
    if fp64:
        f.set_precision(64)
    ew = oracle.Ewald(c.lat, 0.35, gcell=8)
    of = oracle.ion_field(["M", "X"], pots, rc=10.0, ewald=ew, min_image=False, lat=c.lat)
    ref = oracle.ion_forces(of, c.pos, c.types, c.charges, c.lat)
    got = f.calculate_forces(c.pos, c.types, c.lat, c.charges)
    tol = (1e-8, 1e-10) if fp64 else (F_TOL, E_TOL_ATOM)
    _compare(f"rock-salt {forms} fp{64 if fp64 else 32}", ref, got, c.n, *tol)
    assert abs(got["energy"]["vdw"] - ref["e_vdw"]) / c.n < tol[1]
    of.energy_path = 1
    ref_e = oracle.ion_forces(of, c.pos, c.types, c.charges, c.lat)
    e = f.calculate_energy(c.pos, c.types, c.lat, c.charges)
    print(f"  energy path: vdw {e['vdw']:.6f} (oracle {ref_e['e_vdw']:.6f}; force path {ref['e_vdw']:.6f})",
          flush=True)
    assert abs(e["total"] - ref_e["energy"]) / c.n < tol[1]


@pytest.mark.parametrize("fp64", [False, True])
def test_ionic_damped_shifted_coulomb(lib, oracle, fp64):
    """ewald damp: the damped-shifted Coulomb sum on the erfc mesh of
    fieldErfcGen (5000 points, this build's choice of the reference's
    MAXMESH) with Buckingham, explicit images, against the oracle"""
    c = W.rocksalt_cell(seed=5)
    pots = [("M", "X") + W.ROCKSALT_BUCK[("M", "X")], ("X", "X") + W.ROCKSALT_BUCK[("X", "X")],
            ("M", "M") + W.ROCKSALT_BUCK[("M", "M")]]
    f = lib.Field(lib.buck_potentials_text(["M", "X"], [2.0, -2.0], pots, 10.0, noimage=True, ewald_damp=0.3))
    if fp64:
        f.set_precision(64)
    of = oracle.ion_field(["M", "X"], pots, rc=10.0, ewald=None, min_image=False, lat=c.lat)
    oracle.set_damped(of, 0.3, 10.0)
    ref = oracle.ion_forces(of, c.pos, c.types, c.charges, c.lat)
    got = f.calculate_forces(c.pos, c.types, c.lat, c.charges)
    tol = (1e-8, 1e-10) if fp64 else (F_TOL, E_TOL_ATOM)
    _compare(f"rock-salt damped-shifted Coulomb fp{64 if fp64 else 32}", ref, got, c.n, *tol)
    assert got["energy"]["rcp"] == 0.0
    assert abs(got["energy"]["real"] - ref["e_real"]) / c.n < tol[1]
