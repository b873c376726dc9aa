"""Pipelined campaigns: the next run chunk is launched before the searches
that ended are read and refilled, so a start is queued behind a chunk in
flight, the statuses come from the host's copy and the saddle positions are
copied on a side stream.  With proposals fixed per (walker, search number),
every walker runs the same sequence of searches as in the plain loop, only
later, so the records of each walker agree bitwise over the searches both
runs finished.
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


def _proposer(camp, w):
    """the walker's k-th proposal depends on (w, k) only"""
    k = int(camp._starts[w])
    rng = np.random.default_rng([5, w, k])
    c = camp.cells[w]
    v = int(rng.integers(0, len(c.vacancies)))
    frozen, nb = camp._site_region(w, v)
    basin = camp.basins[w]
    disp, mode = W.local_displacement(W.Cell(basin, c.types, c.lat), nb, camp.d0, int(rng.integers(0, 2**31 - 1)))
    return basin + disp, mode, frozen, c.vacancies[v]


def _campaign(lib, cells):
    from akmc_amd import search as S
    camp = S.SaddleCampaign(_text(lib), cells, params=lib.dimer_defaults(maxParIter=400), seed=4,
                            proposer=_proposer)
    camp.relax_basins()
    camp.begin()
    return camp


def _by_walker(records):
    out = {}
    for r in records:
        out.setdefault(r.walker, []).append(r)
    return out


def test_pipelined_campaign_matches_plain_loop(lib):
    from akmc_amd import search as S
    cells = W.batched_sc_cells(n_walkers=512, seed=3)
    cells = [cells[i] for i in np.argsort([c.n for c in cells])[:24]]
    plain = _campaign(lib, cells)
    st_plain = plain.advance(1500, sync_every=25)
    pipe = _campaign(lib, cells)
    st_pipe = S.CampaignStats()
    for _ in range(3):
        st_pipe.add(pipe.advance(500, sync_every=25, pipelined=True))
    assert pipe._inflight
    pipe.drain(st_pipe, refill=False)
    assert not pipe._inflight
    print(f"plain: {st_plain.searches} ended, {st_plain.verified} verified; pipelined: {st_pipe.searches} ended, "
          f"{st_pipe.verified} verified", flush=True)
    assert st_plain.searches >= len(cells) and st_pipe.searches >= len(cells)
    a, b = _by_walker(st_plain.records), _by_walker(st_pipe.records)
    compared = 0
    for w in range(len(cells)):
        for ra, rb in zip(a.get(w, []), b.get(w, [])):
            assert (ra.verified, ra.distinct) == (rb.verified, rb.distinct), w
            for k in ("energy32", "energy64", "barrier32", "barrier64", "curvature", "fmax64"):
                x, y = getattr(ra, k), getattr(rb, k)
                assert x == y or (np.isnan(x) and np.isnan(y)), (w, k)
            assert tuple(ra.lanczos) == tuple(rb.lanczos)
            assert np.array_equal(ra.atoms, rb.atoms) and np.array_equal(ra.disp, rb.disp)
            compared += 1
    print(f"records compared: {compared}", flush=True)
    assert compared >= len(cells)
    # the campaign goes on in the plain loop after a drain
    st = pipe.advance(100, sync_every=25)
    assert st.micro_steps == 100
