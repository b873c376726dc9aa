"""GPU checks of the learned proposal: BF16 MFMA forward and gradient
against the FP64 oracle, device descriptors against the FP64 oracle,
training on the device, and the gain of learned proposals over random
displacements on identical cells.

Tolerances: BF16 operands (8-bit mantissa, relative rounding 2^-9) with
FP32 accumulation over K <= 128 give ~1e-2 relative error on outputs and
gradients; the budget below is 3e-2 (relative, norm-wise per block).
Descriptors are FP32 sums of O(100) terms: 1e-4 relative per feature.
"""
import numpy as np
import pytest

from akmc_amd import workloads as W
from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu
CU = W.SC_CU


@pytest.fixture(scope="module")
def learn():
    from akmc_amd import learn as LN
    return LN


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_mfma_forward_and_gradient_match_oracle(learn):
    rng = np.random.default_rng(4)
    p = MO.init_params(7)
    p["b1"] = rng.normal(0, 0.1, 128)
    p["b2"] = rng.normal(0, 0.1, 128)
    p["b3"] = rng.normal(0, 0.1, 3)
    M = 777  # not a multiple of the 64-row tile
    X = rng.normal(size=(M, 48))
    T = rng.normal(size=(M, 3))
    net = learn.MLP(max_rows=1024)
    net.params = learn.pack(p)
    Y = net.forward(X)
    Yr, _ = MO.forward(p, X)
    e = _rel(Y, Yr)
    print(f"forward rel err {e:.2e}", flush=True)
    assert e < 3e-2
    loss, g = net.loss_grad(X, T)
    lr, gr = MO.loss_and_grad(p, X, T)
    gd = learn.unpack(g)
    print(f"loss {loss:.6f} vs {lr:.6f}", flush=True)
    assert abs(loss - lr) < 3e-2 * lr
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        ek = _rel(gd[k], gr[k])
        print(f"grad {k} rel err {ek:.2e}", flush=True)
        assert ek < 3e-2, k


def test_descriptors_match_oracle(learn):
    from akmc_amd import lib as L
    c = W.plain_sc_workload(seed=1)
    f = L.Field(L.sc_potentials_text([("Cu", "Cu", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"])], 2 * CU["a"]))
    b = L.Batch(f, [c])
    b.forces()
    site = c.vacancies[0]
    nb = W.neighbours_of_site(c, site, 12)
    atoms = list(range(0, c.n, 7))
    mover = c.pos[:, nb[3]]
    X = learn.batch_descriptors(b, atoms, np.tile(site, (len(atoms), 1)), np.tile(mover, (len(atoms), 1)))
    R = MO.descriptors(c.pos, c.lat, atoms, site, mover)
    err = np.abs(X - R) / np.maximum(np.abs(R), 1e-3)
    print(f"descriptor max rel err {err.max():.2e}", flush=True)
    assert err.max() < 1e-4


def test_training_reduces_loss(learn):
    rng = np.random.default_rng(1)
    M = 20000
    X = rng.normal(size=(M, 48)).astype(np.float32)
    A = rng.normal(size=(48, 8)) / 7.0
    B = rng.normal(size=(8, 3))
    T = (np.tanh(X @ A) @ B).astype(np.float32)
    net = learn.MLP(max_rows=4096, seed=3)
    net.fit_scaling(X)
    hist = net.train(X, T, epochs=30, batch=1024, lr=2e-3, seed=5)
    print("loss history", np.round(hist[[0, 9, 19, 29]], 5), flush=True)
    assert hist[-1] < 0.1 * hist[0]


def test_learned_proposals_beat_random_displacements(learn):
    """Train on saddles from random-displacement dimers, then run one search
    per walker from random and from learned proposals on the same relaxed
    cells: the learned starts need fewer force calls per FP64-verified
    saddle."""
    from akmc_amd import lib as L
    from akmc_amd import search as S
    cells = W.batched_sc_cells(n_walkers=24, n_min=500, n_max=1500, seed=21)
    text = L.sc_potentials_text([("Cu", "Cu", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"])], 2 * CU["a"])
    camp = S.SaddleCampaign(text, cells, params=L.dimer_defaults(maxParIter=400), seed=2)
    camp.relax_basins()
    camp.begin()
    train_stats = S.CampaignStats()
    for _ in range(8):
        train_stats.add(camp.advance(250, sync_every=25, refill=True))
    while camp.pending:  # let the running searches end
        camp.harvest(train_stats)
        if camp.pending:
            camp.batch.run(25, sync_every=25)
    X, T = learn.dataset_from_records(camp, [r for r in train_stats.records if r.verified])
    print(f"training set: {train_stats.verified} verified saddles, {X.shape[0]} atom rows", flush=True)
    assert X.shape[0] > 2000
    net = learn.MLP(max_rows=4096, seed=1)
    net.fit_scaling(X)
    hist = net.train(X, T, epochs=40, batch=2048, lr=2e-3, seed=3)
    print("training loss", np.round(hist[[0, 19, 39]], 5), flush=True)
    assert hist[-1] < hist[0]
    res = {}
    for name, prop in (("random", None), ("learned", learn.LearnedProposer(camp, net, seed=4))):
        camp.proposer = prop
        st = camp.run(sync_every=25)
        calls = st.force_calls / max(st.verified, 1)
        res[name] = (st.verified, st.searches, calls)
        print(f"{name}: {st.verified}/{st.searches} verified, {calls:.1f} force calls per verified saddle",
              flush=True)
    assert res["learned"][0] >= 0.5 * res["random"][0]
    assert res["learned"][2] < res["random"][2]
