"""The full-loop workload on the GPU: the 31,980-atom two-species
Sutton-Chen alloy (Cu/Ni, 20 x 20 x 20 FCC, 50/50 seeded occupancy, 20
vacancies, explicit Cu-Cu, Ni-Ni and Cu-Ni lines).

- one force and energy evaluation of the whole cell against the FP64 oracle
  (FP32 pair terms: forces within 1e-5 eV/A, energy within 1e-6 eV/atom;
  FP64 path 1e-10);
- a local search cell (the K = 3600 atoms nearest a vacancy) gives the
  whole cell's forces on its free atoms and the whole cell's energy
  differences (device FP64 cell against the oracle's whole cell);
- two KMC steps of the loop with 64 walkers, re-checked on the CPU: every
  total rate is the Arrhenius sum over the catalogue's barriers, the chosen
  barrier lies in the window, events' barriers are the oracle's whole-cell
  energy differences, and each new basin 1 energy is the oracle's.
"""
import numpy as np
import pytest

from akmc_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from akmc_amd import lib as L
    L.load()
    return L


@pytest.fixture(scope="module")
def alloy():
    return W.alloy_cell(seed=7)


def _oracle_field(oracle):
    return oracle.sc_field(["Cu", "Ni"], W.ALLOY_POTS, W.ALLOY_RC)


def test_alloy_cell_forces(lib, oracle, alloy):
    c = alloy
    assert c.n == 31980 and len(c.vacancies) == 20
    assert abs((c.types == 0).mean() - 0.5) < 0.02
    of = _oracle_field(oracle)
    ref = oracle.sc_forces(of, c.pos, c.types, c.lat, omp=True)
    f = lib.Field(W.alloy_potentials_text())
    got = f.calculate_forces(c.pos, c.types, c.lat)
    df = float(np.abs(got["force"] - ref["force"]).max())
    de = abs(got["energy"]["total"] - ref["energy"]) / c.n
    f.set_precision(64)
    g64 = f.calculate_forces(c.pos, c.types, c.lat)
    df64 = float(np.abs(g64["force"] - ref["force"]).max())
    de64 = abs(g64["energy"]["total"] - ref["energy"]) / c.n
    print(f"alloy 31980 atoms: FP32 max|dF| {df:.2e} eV/A |dE|/N {de:.2e} eV/atom; FP64 {df64:.2e} / {de64:.2e}",
          flush=True)
    assert df < 1e-5 and de < 1e-6
    assert df64 < 1e-10 and de64 < 1e-12


def test_local_search_cell_is_exact(lib, oracle, alloy):
    """forces on the free atoms of a K-atom search cell equal the whole
    cell's, and an energy difference of the search cell equals the whole
    cell's (device FP64 cell, oracle FP64 whole cell)"""
    from akmc_amd import kmc_cluster as KC
    c = alloy
    site = c.vacancies[0]
    idx, d, r_in = KC.nearest_atoms(c.pos, c.lat, site, 3600)[0]
    free = d < 6.0
    print(f"search cell: r_in {r_in:.3f} A, {free.sum()} free atoms", flush=True)
    assert r_in > 6.0 + 2 * W.ALLOY_RC
    of = _oracle_field(oracle)
    f = lib.Field(W.alloy_potentials_text())
    f.set_precision(64)
    f.set_list_density(c.n / (c.lat[0] * c.lat[4] * c.lat[8]))
    sub = W.Cell(c.pos[:, idx], c.types[idx], c.lat)
    b = lib.Batch(f, [sub])
    e0 = b.forces()[0]["total"]
    _, frc = b.download()
    ref = oracle.sc_forces(of, c.pos, c.types, c.lat, omp=True)
    dfree = float(np.abs(frc[0][:, free] - ref["force"][:, idx[free]]).max())
    rng = np.random.default_rng(3)
    disp = np.zeros((3, 3600))
    disp[:, free] = rng.normal(0.0, 0.05, (3, int(free.sum())))
    b.upload([W.Cell(sub.pos + disp, sub.types, c.lat)])
    e1 = b.forces()[0]["total"]
    full = c.pos.copy()
    full[:, idx] += disp
    ref1 = oracle.sc_forces(of, full, c.types, c.lat, omp=True)
    dde = abs((e1 - e0) - (ref1["energy"] - ref["energy"]))
    print(f"free-atom forces vs the whole cell {dfree:.2e} eV/A; energy difference {e1 - e0:.8f} vs "
          f"{ref1['energy'] - ref['energy']:.8f} eV ({dde:.2e})", flush=True)
    assert dfree < 1e-10
    assert dde < 1e-8


def test_cluster_kmc_steps(lib, oracle, alloy):
    from akmc_amd import kmc_cluster as KC
    from akmc_amd.kmc import KMCParams
    from oracle import kmc_oracle as K
    c = alloy
    p = KMCParams(kmcTemperature=600.0, kmcPreFactor=1.0e13)
    run = KC.ClusterKMC(W.alloy_potentials_text(), c, walkers=64, cluster_atoms=3600, params=p, seed=5,
                        learned_fraction=0.5, train_epochs=10)
    st = run.run(2)
    print(f"cluster KMC: {st.kmc_steps} steps, {st.searches} searches ({st.learned_searches} learned, "
          f"{st.fallback_searches} fallback), {st.valid_events} valid events ({st.learned_events} learned), "
          f"{st.outside_cell} outside the cell, {st.cancelled} cancelled, {st.seconds:.1f} s", flush=True)
    assert st.kmc_steps == 2
    for rec in st.steps:
        tot = sum(K.rate(e, p.kmcPreFactor, p.kmcTemperature) for e in rec.barriers)
        assert abs(tot - rec.total_rate) <= 1e-12 * tot
        assert p.kmcMinCap <= rec.activation_energy <= p.kmcMaxCap
    of = _oracle_field(oracle)
    # the last catalogue's barriers are the whole cell's FP64 energy differences
    basin1, events, chosen = run.last_catalogue
    e1 = oracle.sc_forces(of, basin1, c.types, c.lat, omp=True)["energy"]
    picks = sorted({chosen} | set(range(min(2, len(events)))))
    for k in picks:
        ev = events[k]
        sad = basin1.copy()
        sad[:, ev.atoms] = ev.saddle
        es = oracle.sc_forces(of, sad, c.types, c.lat, omp=True)["energy"]
        print(f"event {k} (walker {ev.walker}, learned {ev.learned}): barrier {ev.activation_energy:.6f} eV, "
              f"whole-cell oracle {es - e1:.6f} eV", flush=True)
        assert abs((es - e1) - ev.activation_energy) < 1e-6
    e_final = oracle.sc_forces(of, run.basin, c.types, c.lat, omp=True)["energy"]
    print(f"final basin 1: device {run.e_full:.8f} oracle {e_final:.8f}", flush=True)
    assert abs(run.e_full - e_final) < 1e-9 * abs(e_final)
