"""Tabulated EAM fixtures for the tests: LAMMPS setfl files whose tables
sample the Sutton-Chen functions (F = -sqrt(rho), rho(r) = (eps c)^2
(a/r)^m, phi(r) = eps (a/r)^n), written in the layout readLEAM reads, so
the tabulated field can be checked against the independent Sutton-Chen
oracle as the mesh is refined."""
import math

from oracle import eam_oracle as E


def sc_setfl(path, eps, a, n, m, c, rc, dr=1.0e-3, drho=1.0e-3, rho_max=40.0):
    cut = rc + 0.05                      # the field's own cutoff (rc) decides which pairs count
    nr = int(math.ceil((cut + 2 * dr) / dr)) + 1
    nrho = int(math.ceil(rho_max / drho)) + 1
    rr = [max(i * dr, 0.5) for i in range(nr)]   # the r = 0 end is never evaluated
    F = [-math.sqrt(i * drho) for i in range(nrho)]
    rho = [(eps * c) ** 2 * (a / r) ** m for r in rr]
    rphi = [i * dr * eps * (a / r) ** n for i, r in enumerate(rr)]
    E.write_lammps(path, nrho, drho, nr, dr, cut, F, rho, rphi)
    return path
