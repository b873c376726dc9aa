"""Learned saddle proposals: descriptors + MLP on the device.

A proposal for a walker picks one of its vacancies and a candidate mover
among the vacancy's 12 neighbours.  For every atom of the active region the
network maps the atom's 48 descriptors (computed at the basin, with
vacancy-site and mover channels) to a 3-vector in the hop frame
(h = unit(vacancy - mover), e2, e3); back in the lab frame these form the
predicted saddle displacement.  The search then starts from
basin + ``start_fraction`` x prediction with the prediction as its start
mode.  Training pairs come from FP64-verified saddles found by the plain
random-displacement searches (``dataset_from_records``); the
random-displacement dimer remains both the fallback and the judge, since
every saddle reached from a learned proposal is verified exactly like any
other (akmc_amd/search.py).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L
from . import workloads as W

NPARAM = 23171
_F = C.POINTER(C.c_float)


def _setup(lib):
    lib.akmc_mlp_create.argtypes = [C.c_int, C.c_uint, C.POINTER(C.c_void_p)]
    lib.akmc_mlp_destroy.argtypes = [C.c_void_p]
    lib.akmc_mlp_set_params.argtypes = [C.c_void_p, _F]
    lib.akmc_mlp_get_params.argtypes = [C.c_void_p, _F]
    lib.akmc_mlp_forward.argtypes = [C.c_void_p, _F, C.c_int, _F]
    lib.akmc_mlp_loss_grad.argtypes = [C.c_void_p, _F, _F, C.c_int, C.POINTER(C.c_double), _F]
    lib.akmc_mlp_train.argtypes = [C.c_void_p, _F, _F, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint,
                                   C.POINTER(C.c_double)]
    lib.akmc_mlp_time.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]
    lib.akmc_batch_descriptors.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                           C.POINTER(C.c_double), _F]


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a):
    return a.ctypes.data_as(_F)


def pack(p: dict) -> np.ndarray:
    """W1, b1, W2, b2, W3, b3 -> the library's flat parameter order."""
    return _f32(np.concatenate([np.asarray(p[k], dtype=np.float64).reshape(-1)
                                for k in ("W1", "b1", "W2", "b2", "W3", "b3")]))


def unpack(v) -> dict:
    v = np.asarray(v, dtype=np.float64)
    shapes = [("W1", (48, 128)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)), ("W3", (128, 3)), ("b3", (3,))]
    out, o = {}, 0
    for k, sh in shapes:
        n = int(np.prod(sh))
        out[k] = v[o:o + n].reshape(sh)
        o += n
    return out


class MLP:
    """The device MLP (akmc_mlp_* of include/akmc.h)."""

    def __init__(self, max_rows=1 << 16, seed=0):
        lib = L.load()
        _setup(lib)
        h = C.c_void_p()
        L.check(lib.akmc_mlp_create(int(max_rows), int(seed), C.byref(h)))
        self._h = h
        self.max_rows = max_rows
        self.mean = np.zeros(48, np.float32)
        self.scale = np.ones(48, np.float32)

    def close(self):
        if getattr(self, "_h", None):
            L.load().akmc_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(NPARAM, np.float32)
        L.check(L.load().akmc_mlp_get_params(self._h, _fp(p)))
        return p

    @params.setter
    def params(self, p):
        p = _f32(p)
        assert p.size == NPARAM
        L.check(L.load().akmc_mlp_set_params(self._h, _fp(p)))

    def forward(self, X):
        X = _f32(X)
        Y = np.zeros((X.shape[0], 3), np.float32)
        L.check(L.load().akmc_mlp_forward(self._h, _fp(X), X.shape[0], _fp(Y)))
        return Y

    def loss_grad(self, X, T):
        X, T = _f32(X), _f32(T)
        g = np.zeros(NPARAM, np.float32)
        loss = C.c_double()
        L.check(L.load().akmc_mlp_loss_grad(self._h, _fp(X), _fp(T), X.shape[0], C.byref(loss), _fp(g)))
        return loss.value, g

    def fit_scaling(self, X):
        self.mean = _f32(X.mean(0))
        self.scale = _f32(1.0 / np.maximum(X.std(0), 1e-6))

    def standardize(self, X):
        return (np.asarray(X, np.float32) - self.mean) * self.scale

    def train(self, X, T, epochs=20, batch=4096, lr=1e-3, seed=0):
        """Adam on standardised inputs; returns the per-epoch loss."""
        Xs = _f32(self.standardize(X))
        T = _f32(T)
        hist = np.zeros(epochs)
        L.check(L.load().akmc_mlp_train(self._h, _fp(Xs), _fp(T), Xs.shape[0], int(epochs),
                                        int(min(batch, self.max_rows)), float(lr), int(seed),
                                        hist.ctypes.data_as(C.POINTER(C.c_double))))
        return hist

    def time(self, rows, nrep=20):
        """(ms per forward, ms per forward + backward) over `rows` rows."""
        ms = (C.c_double * 2)()
        L.check(L.load().akmc_mlp_time(self._h, int(rows), int(nrep), ms))
        return ms[0], ms[1]

    def predict(self, X):
        return self.forward(self.standardize(X))


def batch_descriptors(batch: L.Batch, atoms_global, vac, mover):
    """Device descriptors for queries at the batch's resident positions."""
    lib = L.load()
    _setup(lib)
    a = np.ascontiguousarray(atoms_global, dtype=np.int32)
    v = np.ascontiguousarray(vac, dtype=np.float64).reshape(-1)
    m = np.ascontiguousarray(mover, dtype=np.float64).reshape(-1)
    X = np.zeros((a.size, 48), np.float32)
    L.check(lib.akmc_batch_descriptors(batch._h, a.size, a.ctypes.data_as(C.POINTER(C.c_int)),
                                       v.ctypes.data_as(C.POINTER(C.c_double)),
                                       m.ctypes.data_as(C.POINTER(C.c_double)), _fp(X)))
    return X


def hop_frame(h):
    h = np.asarray(h, dtype=np.float64)
    h = h / np.linalg.norm(h)
    ax = np.eye(3)[int(np.argmin(np.abs(h)))]
    e2 = np.cross(h, ax)
    e2 /= np.linalg.norm(e2)
    return np.stack([h, e2, np.cross(h, e2)])


def _basin_state(camp, pool=None):
    """Put the campaign's batch at its basins (pool entry `pool` of every
    walker when the campaign turns basins over) with fresh lists."""
    if pool is None or camp.pool is None:
        cells, basins = camp.cells, camp.basins
    else:
        ent = [camp.pool[w][pool % len(camp.pool[w])] for w in range(camp.W)]
        cells, basins = [e[0] for e in ent], [e[1] for e in ent]
    camp.batch.upload([W.Cell(p, c.types, c.lat) for p, c in zip(basins, cells)])
    camp.batch.forces()
    return cells, basins


def dataset_from_records(camp, records):
    """(X, T) from verified saddle records of a campaign: one row per active
    atom of each saddle; T is that atom's saddle displacement in the hop
    frame of the saddle's mover (the atom that moved most).  Records of
    different pooled basins are described at their own basins.  Call it
    when no search is running: it leaves the batch on the campaign's
    current basins."""
    good = [r for r in records if r.verified and r.atoms.size and r.site is not None]
    groups = sorted({r.pool for r in good}) if camp.pool is not None else [None]
    Xs, Ts = [], []
    for g in groups:
        cells, basins = _basin_state(camp, g)
        qa, qv, qm, T = [], [], [], []
        for r in good:
            if g is not None and r.pool != g:
                continue
            w = r.walker
            c, basin = cells[w], basins[w]
            mover = int(r.atoms[np.argmax(np.sqrt((r.disp ** 2).sum(0)))])
            site = r.site
            F = hop_frame(W.min_image(site - basin[:, mover], c.lat))
            act = r.active
            disp = np.zeros((3, c.n))
            disp[:, act] = r.active_disp
            off = camp.batch.offsets[w]
            for i in act:
                qa.append(off + i)
                qv.append(site)
                qm.append(basin[:, mover])
                T.append(F @ disp[:, i])
        if qa:
            Xs.append(batch_descriptors(camp.batch, qa, np.array(qv), np.array(qm)))
            Ts.append(_f32(np.array(T)))
    if groups != [None]:
        _basin_state(camp)
    if not Xs:
        return np.zeros((0, 48), np.float32), np.zeros((0, 3), np.float32)
    return np.concatenate(Xs), np.concatenate(Ts)


def _active_region(cell, pos, site, radius):
    d = W.min_image(pos - np.asarray(site)[:, None], cell.lat)
    return (np.sqrt((d * d).sum(0)) >= radius).astype(np.int32)


class LearnedProposer:
    """Proposals from the trained MLP for every (walker, vacancy, mover)
    candidate, precomputed once at the basins."""

    def __init__(self, camp, mlp: MLP, start_fraction=0.6, noise=0.02, seed=0):
        self.camp = camp
        self.mlp = mlp
        self.f = start_fraction
        self.noise = noise
        self.rng = np.random.default_rng(seed)
        self.cands = {}
        self._precompute()

    def _precompute(self):
        camp = self.camp
        _basin_state(camp)
        qa, qv, qm, keys = [], [], [], []
        for w, c in enumerate(camp.cells):
            if not camp.basin_ok[w]:
                continue
            basin = camp.basins[w]
            off = camp.batch.offsets[w]
            for vi, site in enumerate(c.vacancies):
                frozen = _active_region(c, basin, site, camp.region_radius)
                act = np.nonzero(frozen == 0)[0]
                nb = W.neighbours_of_site(W.Cell(basin, c.types, c.lat), site, 12)
                for m in nb:
                    keys.append((w, vi, int(m), act, frozen))
                    for i in act:
                        qa.append(off + i)
                        qv.append(site)
                        qm.append(basin[:, m])
        X = batch_descriptors(camp.batch, qa, np.array(qv), np.array(qm))
        Y = self.mlp.predict(X).astype(np.float64)
        o = 0
        for (w, vi, m, act, frozen) in keys:
            c = camp.cells[w]
            site = c.vacancies[vi]
            F = hop_frame(W.min_image(site - camp.basins[w][:, m], c.lat))
            pred = np.zeros((3, c.n))
            pred[:, act] = (Y[o:o + act.size] @ F).T
            o += act.size
            self.cands.setdefault(w, []).append((vi, m, pred, frozen))

    def __call__(self, camp, w):
        vi, m, pred, frozen = self.cands[w][int(self.rng.integers(0, len(self.cands[w])))]
        c = camp.cells[w]
        site = c.vacancies[vi]
        disp = self.f * pred
        if self.noise:
            disp = disp + np.where(frozen == 0, 1.0, 0.0) * self.rng.normal(0.0, self.noise, pred.shape)
        mode = pred if np.linalg.norm(pred) > 0 else disp
        return camp.basins[w] + disp, mode, frozen, site
