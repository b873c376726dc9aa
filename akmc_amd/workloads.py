"""Seeded synthetic inputs for the force-and-saddle-search path.

Every configuration is built from numpy's PCG64 generator with an explicit
seed, so the same seed gives the same atoms, vacancies, jitter and start
displacements here, in the tests and in bench.py.  Positions are SoA arrays
of shape (3, N) in Angstrom, lattices are row vectors (lat[0:3] = a1), the
layout of the reference's Basis/Field (Common/Basis.cpp:356-470).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# Sutton-Chen Cu of the plain-method workload: eps, a, n, m, c
SC_CU = dict(eps=0.012382, a=3.61, n=9.0, m=6.0, c=39.432)

# The two-species alloy of the full-loop workload: Sutton-Chen Cu and Ni
# (Sutton & Chen 1990: eps eV, a A, n, m, c) and an explicit Cu-Ni line
# (geometric-mean eps, mean a and c), FCC at the mean lattice constant
SC_NI = dict(eps=0.015707, a=3.52, n=9.0, m=6.0, c=39.432)
ALLOY_POTS = [
    ("Cu", "Cu", SC_CU["eps"], SC_CU["a"], SC_CU["n"], SC_CU["m"], SC_CU["c"]),
    ("Ni", "Ni", SC_NI["eps"], SC_NI["a"], SC_NI["n"], SC_NI["m"], SC_NI["c"]),
    ("Cu", "Ni", (SC_CU["eps"] * SC_NI["eps"]) ** 0.5, 0.5 * (SC_CU["a"] + SC_NI["a"]), 9.0, 6.0,
     0.5 * (SC_CU["c"] + SC_NI["c"])),
]
ALLOY_A = 0.5 * (SC_CU["a"] + SC_NI["a"])
ALLOY_RC = 2.0 * SC_CU["a"]

# Buckingham parameters of the rock-salt workload (A eV, rho A, C eV A^6)
ROCKSALT_BUCK = {
    ("M", "X"): (1284.38, 0.29969, 0.0),
    ("X", "X"): (22764.0, 0.149, 27.88),
    ("M", "M"): (0.0, 1.0, 0.0),
}


@dataclass
class Cell:
    pos: np.ndarray            # (3, N) float64
    types: np.ndarray          # (N,) int32
    lat: np.ndarray            # (9,) float64, rows are lattice vectors
    species: list = field(default_factory=list)
    charges: np.ndarray | None = None
    vacancies: list = field(default_factory=list)   # removed site coordinates (3,)
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.pos.shape[1]


def fcc_sites(a: float, nx: int, ny: int, nz: int) -> np.ndarray:
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]], dtype=np.float64)
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    sites = (g[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a
    return sites.T.copy()


def rocksalt_sites(a: float, nx: int, ny: int, nz: int):
    cat = fcc_sites(a, nx, ny, nz)
    ani = cat + 0.5 * a * np.array([[1.0], [0.0], [0.0]])
    ani[0] = np.mod(ani[0], a * nx)
    return cat, ani


def min_image(d: np.ndarray, lat: np.ndarray) -> np.ndarray:
    """Minimum-image displacement(s) d (3, ...) for an orthorhombic cell."""
    L = np.array([lat[0], lat[4], lat[8]]).reshape((3,) + (1,) * (d.ndim - 1))
    return d - L * np.rint(d / L)


def sc_cell(nx, ny, nz, n_vac=1, jitter=0.01, seed=0, a=SC_CU["a"], two_species=False):
    """FCC Sutton-Chen cell with vacancies at seeded random sites and
    N(0, jitter) displacements."""
    rng = np.random.default_rng(seed)
    sites = fcc_sites(a, nx, ny, nz)
    ns = sites.shape[1]
    vac = np.sort(rng.choice(ns, size=n_vac, replace=False)) if n_vac else np.zeros(0, dtype=int)
    keep = np.setdiff1d(np.arange(ns), vac)
    pos = sites[:, keep] + rng.normal(0.0, jitter, size=(3, keep.size)) if jitter else sites[:, keep].copy()
    if two_species:
        types = rng.integers(0, 2, size=ns).astype(np.int32)[keep]
        species = ["A", "B"]
    else:
        types = np.zeros(keep.size, dtype=np.int32)
        species = ["Cu"]
    lat = np.diag([a * nx, a * ny, a * nz]).reshape(9).astype(np.float64)
    return Cell(pos=np.ascontiguousarray(pos), types=types, lat=lat, species=species,
                vacancies=[sites[:, v].copy() for v in vac], meta=dict(seed=seed, nx=nx, ny=ny, nz=nz, jitter=jitter))


def site_distances(cell: Cell, site: np.ndarray) -> np.ndarray:
    """minimum-image distance of every atom from site"""
    d = min_image(cell.pos - np.asarray(site)[:, None], cell.lat)
    return np.sqrt((d * d).sum(0))


def neighbours_of_site(cell: Cell, site: np.ndarray, count: int = 12, r: np.ndarray | None = None) -> np.ndarray:
    """The count atoms nearest to site (minimum image), nearest first, ties
    in index order: a stable sort of the atoms within the count-th distance
    (the first count of a stable sort of all of them, without sorting all).
    r: the distances from site_distances, when already at hand."""
    if r is None:
        r = site_distances(cell, site)
    if count >= r.size:
        return np.argsort(r, kind="stable")[:count]
    cut = np.partition(r, count - 1)[count - 1]
    near = np.nonzero(r <= cut)[0]
    return near[np.argsort(r[near], kind="stable")[:count]]


def local_displacement(cell: Cell, atoms: np.ndarray, d0: float, seed: int):
    """Random displacement of the listed atoms: each component is
    (u - 1/2) d0 with the centre of mass of the displaced set removed, the
    form of makeRandomDisplacement (optim/SaddleSearch.cpp:391-414) applied
    to a local set.  Returns (displacement (3, N), unit mode (3, N))."""
    rng = np.random.default_rng(seed)
    u = rng.random((atoms.size, 3))
    d = u - 0.5
    d -= d.mean(0, keepdims=True)
    disp = np.zeros((3, cell.n))
    disp[:, atoms] = (d * d0).T
    mode = np.zeros((3, cell.n))
    mode[:, atoms] = d.T
    mode /= np.linalg.norm(mode)
    return disp, mode


def plain_sc_workload(seed: int = 1):
    """4x4x4 FCC Sutton-Chen Cu, one vacancy, N(0, 0.01 A) jitter."""
    return sc_cell(4, 4, 4, n_vac=1, jitter=0.01, seed=seed)


def rocksalt_cell(nx=4, ny=4, nz=4, a=4.212, jitter=0.02, seed=0, charge=2.0, schottky=True):
    """Rock-salt cell with charges +-charge, one seeded Schottky pair
    (a cation and its nearest anion) removed, N(0, jitter) displacements."""
    rng = np.random.default_rng(seed)
    cat, ani = rocksalt_sites(a, nx, ny, nz)
    pos = np.concatenate([cat, ani], axis=1)
    types = np.concatenate([np.zeros(cat.shape[1], np.int32), np.ones(ani.shape[1], np.int32)])
    lat = np.diag([a * nx, a * ny, a * nz]).reshape(9).astype(np.float64)
    vac = []
    if schottky:
        ic = int(rng.integers(0, cat.shape[1]))
        d = min_image(ani - cat[:, ic:ic + 1], lat)
        ja = int(np.argmin((d * d).sum(0)))
        remove = [ic, cat.shape[1] + ja]
        vac = [pos[:, ic].copy(), pos[:, cat.shape[1] + ja].copy()]
        keep = np.setdiff1d(np.arange(pos.shape[1]), remove)
        pos, types = pos[:, keep], types[keep]
    if jitter:
        pos = pos + rng.normal(0.0, jitter, size=pos.shape)
    charges = np.where(types == 0, charge, -charge).astype(np.float64)
    return Cell(pos=np.ascontiguousarray(pos), types=types.astype(np.int32), lat=lat, species=["M", "X"],
                charges=charges, vacancies=vac, meta=dict(seed=seed, a=a))


def alloy_cell(n=20, n_vac=20, jitter=0.01, seed=7, frac_a=0.5):
    """The full-loop workload's cell: FCC n x n x n at ALLOY_A, species 0
    (Cu) and 1 (Ni) drawn 50/50 per site from the seed, n_vac vacancies at
    seeded sites, N(0, jitter) displacements (20 x 20 x 20, 20 vacancies:
    31,980 atoms)."""
    rng = np.random.default_rng(seed)
    sites = fcc_sites(ALLOY_A, n, n, n)
    ns = sites.shape[1]
    kind = (rng.random(ns) >= frac_a).astype(np.int32)
    vac = np.sort(rng.choice(ns, size=n_vac, replace=False))
    keep = np.setdiff1d(np.arange(ns), vac)
    pos = sites[:, keep] + rng.normal(0.0, jitter, size=(3, keep.size))
    lat = np.diag([ALLOY_A * n] * 3).reshape(9).astype(np.float64)
    return Cell(pos=np.ascontiguousarray(pos), types=kind[keep], lat=lat, species=["Cu", "Ni"],
                vacancies=[sites[:, v].copy() for v in vac], meta=dict(seed=seed, nx=n, ny=n, nz=n, jitter=jitter))


def alloy_potentials_text():
    from .lib import sc_potentials_text
    return sc_potentials_text(ALLOY_POTS, ALLOY_RC, species=["Cu", "Ni"])


def batched_sc_cells(n_walkers=512, n_min=500, n_max=4000, seed=3, jitter=0.01):
    """Walker cells of the batched workload: FCC n_x x n_y x n_z (all >= 4)
    with 4 n_x n_y n_z in [n_min, n_max] drawn uniformly over the feasible
    shapes, 1-3 vacancies each at seeded sites."""
    rng = np.random.default_rng(seed)
    shapes = [(i, j, k) for i in range(4, 17) for j in range(i, 17) for k in range(j, 17)
              if n_min <= 4 * i * j * k <= n_max]
    cells = []
    for w in range(n_walkers):
        s = shapes[int(rng.integers(0, len(shapes)))]
        nv = int(rng.integers(1, 4))
        cells.append(sc_cell(*s, n_vac=nv, jitter=jitter, seed=int(rng.integers(0, 2**31 - 1))))
    return cells


def write_sc_setfl(path, eps, a, n, m, c, rc, dr=1.0e-3, drho=1.0e-3, rho_max=40.0):
    """A LAMMPS setfl file (the layout of the reference's readLEAM) whose
    tables sample the Sutton-Chen functions: F(rho) = -sqrt(rho),
    rho(r) = (eps c)^2 (a/r)^m and r phi(r) = r eps (a/r)^n, with the r = 0
    end held at r = 0.5 A (never evaluated).  The table cutoff sits 0.05 A
    past rc so the field's cutoff decides which pairs count."""
    import math
    cut = rc + 0.05
    nr = int(math.ceil((cut + 2 * dr) / dr)) + 1
    nrho = int(math.ceil(rho_max / drho)) + 1
    rr = [max(i * dr, 0.5) for i in range(nr)]
    cols = [[-math.sqrt(i * drho) for i in range(nrho)],
            [(eps * c) ** 2 * (a / r) ** m for r in rr],
            [i * dr * eps * (a / r) ** n for i, r in enumerate(rr)]]
    with open(path, "w") as fh:
        fh.write("# Sutton-Chen functions on a mesh\n#\n#\n1 Cu\n")
        fh.write(f"{nrho} {drho!r} {nr} {dr!r} {cut!r}\n29 63.546 {a!r} fcc\n")
        for col in cols:
            for k in range(0, len(col), 5):
                fh.write(" ".join(repr(float(v)) for v in col[k:k + 5]) + "\n")
    return path
