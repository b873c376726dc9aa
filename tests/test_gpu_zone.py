"""Active zones of the search micro-steps: a search evaluates forces on its
free atoms and densities and energies on the atoms within the list radius
of a free atom, and every other atom keeps its values from the walker's
full evaluation at the start.  The bar is bitwise: the same searches with
zones on and off give identical statuses, energies, curvatures, saddle and
basin-2 positions, free-atom forces and campaign records, while the stale
forces on frozen atoms near the free region show that the zones ran.
"""
import numpy as np
import pytest

from akmc_amd import workloads as W

pytestmark = pytest.mark.gpu
CU = W.SC_CU


@pytest.fixture(scope="module")
def lib():
    from akmc_amd import lib as L
    L.load()
    return L


def _text(lib):
    return lib.sc_potentials_text([("Cu", "Cu", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"])], 2 * CU["a"])


def _cells():
    cells = W.batched_sc_cells(n_walkers=512, seed=3)
    order = np.argsort([c.n for c in cells])
    # the smallest cells and a few of 1500-4000 atoms (zones much smaller than the cell)
    return [cells[i] for i in list(order[:3]) + list(order[200:203]) + list(order[-2:])]


def _starts(cells, basins, seed):
    frozen, modes, x0 = [], [], []
    for w, (c, basin) in enumerate(zip(cells, basins)):
        site = c.vacancies[0]
        dv = W.min_image(basin - site[:, None], c.lat)
        fr = (np.sqrt((dv * dv).sum(0)) >= 6.0).astype(np.int32)
        nb = W.neighbours_of_site(W.Cell(basin, c.types, c.lat), site, 12)
        disp, mode = W.local_displacement(W.Cell(basin, c.types, c.lat), nb, 0.1, seed + w)
        frozen.append(fr)
        modes.append(mode)
        x0.append(basin + disp)
    return frozen, modes, x0


def _run(lib, cells, zone):
    """relax, then dimer + FP64 verification + basin 2 on every walker"""
    f = lib.Field(_text(lib))
    b = lib.Batch(f, cells)
    b.set_search_zone(zone)
    b.set_frozen([np.zeros(c.n, np.int32) for c in cells])
    b.start_relax(lib.relax_defaults())
    b.run(20000, sync_every=50)
    basins = b.positions()
    frozen, modes, x0 = _starts(cells, basins, 11)
    b.set_basins()
    b.set_basin2(lib.basin2_defaults())
    b.set_verify(16, 0.01)
    b.start_dimer(modes, params=lib.dimer_defaults(maxParIter=400), x0_per_walker=x0, frozen_per_walker=frozen)
    b.run(40000, sync_every=25)
    out = dict(status=b.status_array(), pos=b.positions(), saddles=b.walkers_saddles(list(range(len(cells)))))
    out["forces"] = b.download()[1]
    out["frozen"] = frozen
    return out


def test_zone_searches_equal_full_bitwise(lib):
    cells = _cells()
    on = _run(lib, cells, True)
    off = _run(lib, cells, False)
    st_on, st_off = on["status"], off["status"]
    print("status", st_on["status"].tolist(), "verified", st_on["verified"].tolist(), "force calls",
          st_on["force_calls"].tolist(), flush=True)
    for k in st_on.dtype.names:
        assert np.array_equal(st_on[k], st_off[k], equal_nan=True), k
    assert (st_on["done"] == 1).all()
    assert (st_on["verified"] == 1).sum() >= len(cells) // 2
    stale = 0
    for w in range(len(cells)):
        assert np.array_equal(on["pos"][w], off["pos"][w])
        assert np.array_equal(on["saddles"][w], off["saddles"][w])
        free = on["frozen"][w] == 0
        assert np.array_equal(on["forces"][w][:, free], off["forces"][w][:, free])
        d = np.abs(on["forces"][w][:, ~free] - off["forces"][w][:, ~free]).max()
        stale += d > 1e-6
    print(f"walkers whose frozen-atom forces were left from the start: {stale} of {len(cells)}", flush=True)
    assert stale >= len(cells) // 2


def test_zone_campaign_records_equal_full(lib):
    """A campaign with basin turnover (restarts on new pool basins, so the
    frozen atoms move between searches) gives the same records with zones
    on and off."""
    from akmc_amd import search as S
    cells = W.batched_sc_cells(n_walkers=512, seed=3)
    cells = [cells[i] for i in np.argsort([c.n for c in cells])[:8]]
    recs = []
    for zone in (True, False):
        camp = S.SaddleCampaign(_text(lib), cells, params=lib.dimer_defaults(maxParIter=400), seed=9)
        camp.batch.set_search_zone(zone)
        camp.relax_basins()
        camp.relax_pool(3, seed=21)
        camp.searches_per_basin = 1
        camp.begin()
        st = camp.advance(1200, sync_every=25)
        recs.append(st.records)
        print(f"zone {zone}: {st.searches} ended, {st.verified} verified, {camp.basin_moves} basin moves", flush=True)
    a, b = recs
    assert len(a) == len(b) and len(a) >= 8
    for ra, rb in zip(a, b):
        assert (ra.walker, ra.pool, ra.verified) == (rb.walker, rb.pool, rb.verified)
        for k in ("energy32", "energy64", "barrier32", "barrier64", "curvature", "fmax64"):
            assert getattr(ra, k) == getattr(rb, k) or (np.isnan(getattr(ra, k)) and np.isnan(getattr(rb, k))), k
        assert tuple(ra.lanczos) == tuple(rb.lanczos)
        assert np.array_equal(ra.disp, rb.disp)
