"""FP64 numpy oracle of the learned proposal (TEST INFRASTRUCTURE ONLY).

Restates, in float64, what the device computes in akmc_amd/csrc/mlp_kernels.hip:
per-atom symmetry-function descriptors (Behler & Parrinello 2007; Behler
2011) and the MLP 48 -> 128 -> 128 -> 3 with SiLU, its mean-squared loss and
the loss gradient by reverse-mode differentiation.  The reference has no
learned component, so this oracle is pinned by its own checks in
tests/test_mlp_oracle.py (finite differences of the loss, descriptor
invariances and closed forms).

Descriptor vector of atom i (48 values): 32 radial G2 and 16 angular G4
symmetry functions.  The radial ones are resolved by neighbour kind, the
atoms of the cell and two pseudo-neighbours that carry the proposal's
context: the vacancy site v and the candidate mover m (the way
element-resolved G2 sets treat a second species).  Cosine cutoff
fc(r) = (cos(pi r/rc)+1)/2 at RC = 7.22 A:
  0..15  G2 over atoms:     sum_j exp(-ETA_R (r_ij - R_k)^2) fc(r_ij),
         R_k = 2.0 + 5.0 k / 15
  16..23 G2 over the site:  exp(-ETA_S (|r_i - v| - S_k)^2) fc(|r_i - v|),
         S_k = 1.5 + 5.0 k / 7
  24..31 G2 over the mover: exp(-ETA_S (|r_i - r_m| - M_k)^2) fc(|r_i - r_m|),
         M_k = 6.0 k / 7
  32..47 G4 over neighbour pairs j < k inside RC_ANG:
         2^(1-zeta) sum (1 + lam cos t_ijk)^zeta exp(-eta (r_ij^2 + r_ik^2 + r_jk^2))
         fa(r_ij) fa(r_ik) fa(r_jk), fa the cosine cutoff at RC_ANG = 5.0,
         (zeta, lam, eta) over zeta in (1, 2, 4, 16), lam in (1, -1),
         eta in (0.005, 0.02), in that nesting order
"""
from __future__ import annotations

import numpy as np

RC = 7.22
RC_ANG = 5.0
ETA_R = 8.0
R_CENTRES = 2.0 + 5.0 * np.arange(16) / 15.0
ETA_S = 2.0
S_CENTRES = 1.5 + 5.0 * np.arange(8) / 7.0
M_CENTRES = 6.0 * np.arange(8) / 7.0
G4_PARAMS = [(z, l, e) for z in (1.0, 2.0, 4.0, 16.0) for l in (1.0, -1.0) for e in (0.005, 0.02)]
NDESC = 48
HIDDEN = 128
NOUT = 3


def fc(r, rc=RC):
    return np.where(r < rc, 0.5 * (np.cos(np.pi * r / rc) + 1.0), 0.0)


def min_image(d, L):
    return d - L * np.rint(d / L)


def descriptors(pos, lat, atoms, vac, mover, rc=RC):
    """pos (3, N); lat 9; atoms: indices to describe; vac, mover: (3,)."""
    L = np.array([lat[0], lat[4], lat[8]])
    out = np.zeros((len(atoms), NDESC))
    P = pos.T
    for a, i in enumerate(atoms):
        d = min_image(P - P[i], L)
        r = np.sqrt((d * d).sum(1))
        nb = np.nonzero((r > 1e-8) & (r < rc))[0]
        rn, dn = r[nb], d[nb]
        f = fc(rn, rc)
        out[a, :16] = (np.exp(-ETA_R * (rn[:, None] - R_CENTRES[None, :]) ** 2) * f[:, None]).sum(0)
        dv = np.linalg.norm(min_image(P[i] - vac, L))
        out[a, 16:24] = np.exp(-ETA_S * (dv - S_CENTRES) ** 2) * fc(dv, rc)
        dm = np.linalg.norm(min_image(P[i] - mover, L))
        out[a, 24:32] = np.exp(-ETA_S * (dm - M_CENTRES) ** 2) * fc(dm, rc)
        an = np.nonzero(rn < RC_ANG)[0]
        if len(an) >= 2:
            ra, da = rn[an], dn[an]
            fa = fc(ra, RC_ANG)
            jj, kk = np.triu_indices(len(an), 1)
            rij, rik = ra[jj], ra[kk]
            rjk_v = da[kk] - da[jj]
            rjk = np.sqrt((rjk_v * rjk_v).sum(1))
            cos = (da[jj] * da[kk]).sum(1) / (rij * rik)
            base = fa[jj] * fa[kk] * fc(rjk, RC_ANG)
            r2 = rij * rij + rik * rik + rjk * rjk
            for t, (z, l, e) in enumerate(G4_PARAMS):
                out[a, 32 + t] = 2.0 ** (1.0 - z) * ((1.0 + l * cos) ** z * np.exp(-e * r2) * base).sum()
    return out


def hop_frame(h):
    """Orthonormal frame (h, e2, e3) of a hop vector h: e2 from the lab axis
    least aligned with h."""
    h = np.asarray(h, dtype=np.float64)
    h = h / np.linalg.norm(h)
    ax = np.eye(3)[int(np.argmin(np.abs(h)))]
    e2 = np.cross(h, ax)
    e2 /= np.linalg.norm(e2)
    e3 = np.cross(h, e2)
    return np.stack([h, e2, e3])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    return {
        "W1": rng.normal(0, np.sqrt(2.0 / NDESC), (NDESC, HIDDEN)), "b1": np.zeros(HIDDEN),
        "W2": rng.normal(0, np.sqrt(2.0 / HIDDEN), (HIDDEN, HIDDEN)), "b2": np.zeros(HIDDEN),
        "W3": rng.normal(0, np.sqrt(1.0 / HIDDEN), (HIDDEN, NOUT)), "b3": np.zeros(NOUT),
    }


def forward(p, X):
    Z1 = X @ p["W1"] + p["b1"]
    H1 = silu(Z1)
    Z2 = H1 @ p["W2"] + p["b2"]
    H2 = silu(Z2)
    Y = H2 @ p["W3"] + p["b3"]
    return Y, (X, Z1, H1, Z2, H2)


def loss_and_grad(p, X, T):
    """L = mean over samples of |Y - T|^2 and dL/dparams."""
    Y, (X, Z1, H1, Z2, H2) = forward(p, X)
    M = X.shape[0]
    R = Y - T
    L = float((R * R).sum() / M)
    dY = 2.0 * R / M
    g = {"W3": H2.T @ dY, "b3": dY.sum(0)}
    dZ2 = (dY @ p["W3"].T) * dsilu(Z2)
    g["W2"] = H1.T @ dZ2
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2 @ p["W2"].T) * dsilu(Z1)
    g["W1"] = X.T @ dZ1
    g["b1"] = dZ1.sum(0)
    return L, g
