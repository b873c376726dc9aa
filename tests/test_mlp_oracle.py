"""Pins of the FP64 learned-proposal oracle (no GPU).

The reference has no learned component; these checks fix the oracle:
loss gradient against central finite differences, descriptor closed forms
for two- and three-atom clusters, invariance under translation and
rotation, and the hop frame's orthonormality.
"""
import numpy as np

from oracle import mlp_oracle as M


def _box(L=60.0):
    return np.diag([L, L, L]).reshape(9)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(1)
    X = rng.normal(size=(37, M.NDESC))
    T = rng.normal(size=(37, M.NOUT))
    L0, g = M.loss_and_grad(p, X, T)
    h = 1e-6
    for name in ("W1", "b1", "W2", "b2", "W3", "b3"):
        flat = p[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp, _ = M.loss_and_grad(p, X, T)
            flat[idx] = old - h
            lm, _ = M.loss_and_grad(p, X, T)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[name].reshape(-1)[idx]) < 1e-7 * max(1.0, abs(fd)), (name, idx)


def test_descriptor_closed_forms():
    L = _box()
    r = 3.1
    pos = np.array([[10.0, 10.0 + r], [10.0, 10.0], [10.0, 10.0]])
    vac = np.array([40.0, 40.0, 40.0])
    d = M.descriptors(pos, L, [0], vac, pos[:, 1])[0]
    f = 0.5 * (np.cos(np.pi * r / M.RC) + 1)
    assert np.allclose(d[:16], np.exp(-M.ETA_R * (r - M.R_CENTRES) ** 2) * f, rtol=1e-13, atol=0)
    # the mover 3.1 A away is a pseudo-neighbour; the vacancy 52 A away is
    # past the cutoff
    assert np.allclose(d[24:32], np.exp(-M.ETA_S * (r - M.M_CENTRES) ** 2) * f, rtol=1e-13)
    assert np.all(d[16:24] == 0.0)
    dv = M.descriptors(pos, L, [0], pos[:, 0] + np.array([0.0, 0.0, 2.2]), pos[:, 1])[0]
    fv = 0.5 * (np.cos(np.pi * 2.2 / M.RC) + 1)
    assert np.allclose(dv[16:24], np.exp(-M.ETA_S * (2.2 - M.S_CENTRES) ** 2) * fv, rtol=1e-13)
    assert np.all(d[32:] == 0.0)  # a single neighbour has no angle
    # equilateral triangle: cos = 1/2 at the apex
    a = 2.7
    tri = np.array([[0.0, a, a / 2], [0.0, 0.0, a * np.sqrt(3) / 2], [0.0, 0.0, 0.0]]) + 20.0
    d = M.descriptors(tri, L, [0], vac, tri[:, 1])[0]
    f = 0.5 * (np.cos(np.pi * a / M.RC_ANG) + 1)
    for t, (z, l, e) in enumerate(M.G4_PARAMS):
        ref = 2.0 ** (1 - z) * (1 + l * 0.5) ** z * np.exp(-e * 3 * a * a) * f ** 3
        assert abs(d[32 + t] - ref) < 1e-14 * max(1.0, ref)


def test_descriptor_invariances():
    rng = np.random.default_rng(3)
    L = _box()
    pos = rng.uniform(25, 35, size=(3, 30))
    vac, mover = np.array([30.0, 30.0, 30.0]), pos[:, 4].copy()
    d0 = M.descriptors(pos, L, range(30), vac, mover)
    sh = np.array([[1.3], [-2.2], [0.7]])
    d1 = M.descriptors(pos + sh, L, range(30), vac + sh[:, 0], mover + sh[:, 0])
    assert np.abs(d0 - d1).max() < 1e-9
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    c = np.array([[30.0], [30.0], [30.0]])
    d2 = M.descriptors(Q @ (pos - c) + c, L, range(30), Q @ (vac - c[:, 0]) + c[:, 0], Q @ (mover - c[:, 0]) + c[:, 0])
    assert np.abs(d0 - d2).max() < 1e-9


def test_hop_frame():
    F = M.hop_frame([0.3, -1.0, 0.2])
    assert np.abs(F @ F.T - np.eye(3)).max() < 1e-14
    assert np.allclose(F[0], np.array([0.3, -1.0, 0.2]) / np.linalg.norm([0.3, -1.0, 0.2]))
