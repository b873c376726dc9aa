"""CPU restatement of the reference's tabulated EAM (LAMMPS setfl tables).

TEST INFRASTRUCTURE ONLY: the checker the device EAM kernels are compared
with.  Nothing in the product path imports it.

Follows potentials/metal/MetalPotential.cpp (readLEAM :276-361,
readParameters :363-386, calculateDerivativeMesh :388-392),
potentials/metal/ManyBody.cpp (eamMeshEnergy :455-460, the eam branches
of calculateManyBodyDensEnergy :539-549, embed :598-615,
calculateManyBodyPairForce :700-707, calculateManyBodyForce :790-825,
calculateManyBodyPairEnergy :444-450, the table of calculateEAMEnergyMesh
:937-1000) and the metal Field's two passes (Field.cpp:293-490 for forces,
:1088-1186 for the energy path).  Every interpolation keeps the
reference's operation order, t_i + ((t_{i+1} - t_i) (x - x_i)) / h with
x_i = i h.  Positions are (3, n) arrays, lattices orthorhombic row-major
lat[9].  Pinned by tests/test_eam_oracle.py: the reader on hand-written
files, exact interpolation of linear tables (energy closed form, forces
= -grad E), perfect-lattice symmetry, and convergence to the Sutton-Chen
oracle for tables sampled from the Sutton-Chen functions.
"""
from __future__ import annotations

import numpy as np

TINY = 1.0e-8


class Table:
    """One MetalPotential as readLEAM leaves it."""

    def __init__(self, nrho, drho, nr, dr, cut, F, rho, rphi):
        self.nrho, self.drho, self.nr, self.dr, self.cut = int(nrho), float(drho), int(nr), float(dr), float(cut)
        self.F = np.asarray(F, dtype=np.float64)
        self.rho = np.asarray(rho, dtype=np.float64)
        with np.errstate(divide="ignore", invalid="ignore"):
            self.phi = np.asarray(rphi, dtype=np.float64) / (np.arange(self.nr) * self.dr)
        self.dF = self._deriv(self.F, self.drho)
        self.drho_t = self._deriv(self.rho, self.dr)
        self.dphi = self._deriv(self.phi, self.dr)

    @staticmethod
    def _deriv(t, h):
        """calculateDerivativeMesh: forward differences, last entry 0"""
        d = np.zeros_like(t)
        with np.errstate(invalid="ignore"):
            d[:-1] = (t[1:] - t[:-1]) / h
        return d


def _read_values(lines, at, n):
    """readParameters: whole lines until n values are in (extra numbers on
    the last line are dropped)"""
    out = []
    while len(out) < n:
        for w in lines[at].split():
            if len(out) < n:
                out.append(float(w))
        at += 1
    return out, at


def read_lammps(path_or_text, is_text=False) -> Table:
    """readLEAM: four header lines, 'Nrho drho Nr dr cutoff', one element
    line, F on the rho mesh, rho and r*phi on the r mesh."""
    text = path_or_text if is_text else open(path_or_text).read()
    lines = text.splitlines()
    w = lines[4].split()
    nrho, drho, nr, dr, cut = int(w[0]), float(w[1]), int(w[2]), float(w[3]), float(w[4])
    at = 6
    F, at = _read_values(lines, at, nrho)
    rho, at = _read_values(lines, at, nr)
    rphi, at = _read_values(lines, at, nr)
    return Table(nrho, drho, nr, dr, cut, F, rho, rphi)


def interp(t, x, h):
    """eamMeshEnergy at index int(x / h) (vectorised; x >= 0)"""
    x = np.asarray(x, dtype=np.float64)
    i = (x / h).astype(np.int64)
    i = np.minimum(i, t.size - 2)
    return t[i] + ((t[i + 1] - t[i]) * (x - i * h)) / h


def _rho_index_ok(tab, rho):
    i = np.floor(np.asarray(rho) / tab.drho)
    return bool(np.all((i >= 0) & (i < tab.nrho)))


class Field:
    """A tabulated EAM field: species names, (A, B, Table) entries, the
    field cutoff rc (rsq <= rc^2, as the potentials text's cutoff)."""

    def __init__(self, species, entries, rc, min_image=True):
        self.species = list(species)
        ns = len(self.species)
        self.tables = [t for _, _, t in entries]
        self.tab = -np.ones((ns, ns), dtype=np.int64)
        for k, (A, B, _) in enumerate(entries):
            for i in range(ns):
                for j in range(ns):
                    if self.species[i] == A and self.species[j] == B:
                        self.tab[j, i] = k
                        self.tab[i, j] = k
        self.rc = float(rc)
        self.min_image = min_image


def _pairs(f: Field, pos, lat, a):
    """displacements r_a - r_b and indices b of the pairs of atom a inside
    rsq <= rc^2 (minimum image, or explicit images with rsq > 1e-3)"""
    L = np.array([lat[0], lat[4], lat[8]])
    n = pos.shape[1]
    d = pos[:, a:a + 1] - pos
    if f.min_image:
        d = d - L[:, None] * np.rint(d / L[:, None])
        idx = np.arange(n)
        keep = (idx != a) & ((d * d).sum(0) <= f.rc * f.rc)
        return d[:, keep], idx[keep]
    cells = [int(f.rc / L[k]) + 1 for k in range(3)]
    ds, js = [], []
    for nx in range(-cells[0], cells[0] + 1):
        for ny in range(-cells[1], cells[1] + 1):
            for nz in range(-cells[2], cells[2] + 1):
                dd = d - np.array([[nx * L[0]], [ny * L[1]], [nz * L[2]]])
                r2 = (dd * dd).sum(0)
                keep = (r2 <= f.rc * f.rc) & (r2 > 1.0e-3)
                ds.append(dd[:, keep])
                js.append(np.nonzero(keep)[0])
    return np.concatenate(ds, axis=1), np.concatenate(js)


def _wrap(f, pos, lat):
    L = np.array([lat[0], lat[4], lat[8]])[:, None]
    s = pos / L
    if f.min_image:
        s = s - np.rint(s)
    else:
        s = s - np.floor(s)
    return s * L


def forces(f: Field, pos, types, lat):
    """calculateForces of the metal Field with eam tables: returns dict
    (energy, many, pair, force (3, n), rho (n,), stress xx yy zz yz xz xy);
    positions wrapped first."""
    pos = _wrap(f, np.asarray(pos, dtype=np.float64), lat)
    types = np.asarray(types)
    n = pos.shape[1]
    rho = np.zeros(n)
    pairs = [_pairs(f, pos, lat, a) for a in range(n)]
    for a in range(n):
        d, js = pairs[a]
        r = np.sqrt((d * d).sum(0))
        for j, rr in zip(js, r):
            k = f.tab[types[j], types[j]]
            if k < 0 or rr < TINY:
                continue
            t = f.tables[k]
            if rr < t.cut:
                rho[a] += interp(t.rho, rr, t.dr)
    many = 0.0
    for a in range(n):
        k = f.tab[types[a], types[a]]
        if k >= 0:
            t = f.tables[k]
            if not _rho_index_ok(t, rho[a]):
                raise ValueError("the rho for embedding is out of bounds")
            many += float(interp(t.F, rho[a], t.drho))
    frc = np.zeros((3, n))
    stress = np.zeros(6)
    epair = 0.0
    for a in range(n):
        d, js = pairs[a]
        r2 = (d * d).sum(0)
        r = np.sqrt(r2)
        for c, j in enumerate(js):
            k = f.tab[types[a], types[j]]
            if k < 0 or r[c] < TINY:
                continue
            t = f.tables[k]
            if not r[c] < t.cut:
                continue
            es = float(interp(t.phi, r[c], t.dr))
            fs = -r[c] * float(interp(t.dphi, r[c], t.dr))
            fm = 0.0
            if rho[a] >= TINY and rho[j] >= TINY:
                fa = float(interp(t.dF, rho[a], t.drho))
                fb = float(interp(t.dF, rho[j], t.drho))
                fm = ((fa + fb) * float(interp(t.drho_t, r[c], t.dr))) * r[c]
            ft = (fs - fm) / (r2[c] * 2.0)
            epair += 0.5 * es
            frc[:, a] += d[:, c] * ft
            frc[:, j] -= d[:, c] * ft
            x, y, z = d[:, c]
            stress -= ft * np.array([x * x, y * y, z * z, y * z, x * z, x * y])
    return dict(energy=many + epair, many=many, pair=epair, force=frc, rho=rho, pos=pos, stress=stress)


def energy_path(f: Field, pos, types, lat):
    """calculateEnergy of the metal Field (Field.cpp:1088-1186): the pair
    term of an eam table is doubled there (ManyBody.cpp:444-450)."""
    r = forces(f, pos, types, lat)
    return r["many"] + 2.0 * r["pair"]


def write_lammps(path, nrho, drho, nr, dr, cut, F, rho, rphi, per_line=5, element="Cu 29 63.546 3.61 fcc"):
    """A setfl file in the layout readLEAM reads (for test fixtures)."""
    def block(v):
        v = list(v)
        return "\n".join(" ".join(repr(float(x)) for x in v[k:k + per_line]) for k in range(0, len(v), per_line))
    with open(path, "w") as fh:
        fh.write("# tabulated test potential\n# \n# \n1 Cu\n")
        fh.write(f"{nrho} {drho!r} {nr} {dr!r} {cut!r}\n{element}\n")
        fh.write(block(F) + "\n" + block(rho) + "\n" + block(rphi) + "\n")
