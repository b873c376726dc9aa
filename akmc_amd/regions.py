"""Active regions and start displacements of the reference's saddle
searches, for the batched engine.

SaddleSearch::constructAtoms (optim/SaddleSearch.cpp:26-349) picks the free
atoms of a search:
  style 1  every atom not frozen already; with a Gaussian start, a random
           free atom is the Gaussian's centre (:41-74)
  style 2  a random free atom and every atom within regionRadius of it
           (:76-138)
  style 4  the same around a random free atom of a species drawn from
           artTypes1 (:140-231)
  style 5  the same around a random free atom of a species drawn from
           artTypes1 whose coordination number (neighbours of the same
           list entry within coordDistance, Basis::calculateCoordNumber,
           Common/Basis.cpp:1300-1360) differs from idealCoordNumber
           (:233-349)
and makeRandomDisplacement / makeGaussDisplacement (:355-491) displace the
free atoms by (u - 1/2) d0, times exp(-d^2 / (2 w^2)) for the Gaussian
start, with the centre of mass of the displacement removed; the start
direction is the normalised unweighted displacement.

Random numbers come from a callable returning uniforms in [0, 1), drawn in
the reference's order, so a run is replayable and the oracle can be fed the
same draws.  Cells are orthorhombic (minimum image per axis).
"""
from __future__ import annotations

import numpy as np

from . import workloads as W


def coordination_numbers(pos, types, lat, typ1, typ2, dist):
    """Basis::calculateCoordNumber: for atoms of type typ1, the number of
    typ2 atoms (other than itself) within dist; 0 for other atoms."""
    pos = np.asarray(pos, dtype=np.float64)
    types = np.asarray(types)
    n = pos.shape[1]
    out = np.zeros(n, dtype=np.int64)
    idx2 = np.nonzero(types == typ2)[0]
    for i in np.nonzero(types == typ1)[0]:
        d = W.min_image(pos[:, idx2] - pos[:, i:i + 1], lat)
        r2 = (d * d).sum(0)
        out[i] = int(np.count_nonzero((r2 <= dist * dist) & (idx2 != i)))
    return out


def _random_atom(uniform, n, ok):
    """the rejection loop of constructAtoms: j = int(u n) until ok(j)"""
    if not np.any([ok(j) for j in range(n)]):
        raise ValueError("no atom qualifies as the centre of the active region")
    while True:
        j = int(uniform() * n)
        if ok(j):
            return j


def choose_centre(style, pos, types, frozen, lat, uniform, species=None, art_types=None, coord_distance=None,
                  ideal_coord=None):
    """The centre atom of region styles 1 (Gaussian start), 2, 4 and 5, with
    the reference's draws (one uniform for the list entry of styles 4 and
    5, then the rejection loop over atoms)."""
    n = len(frozen)
    if style in (1, 2):
        return _random_atom(uniform, n, lambda j: frozen[j] == 0)
    if style not in (4, 5):
        raise ValueError(f"region style {style} is not one of 1, 2, 4, 5")
    choice = int(len(art_types) * uniform())
    name = art_types[choice]
    if name not in species:
        raise ValueError("the atom type has not been found in the species list")
    typ = species.index(name)
    if style == 4:
        return _random_atom(uniform, n, lambda j: types[j] == typ and frozen[j] == 0)
    # style 5: both types come from artTypes1 (SaddleSearch.cpp:263-276)
    cn = coordination_numbers(pos, types, lat, typ, typ, coord_distance[choice])
    return _random_atom(uniform, n, lambda j: types[j] == typ and frozen[j] == 0 and cn[j] != ideal_coord[choice])


def sphere_region(pos, lat, centre_atom, radius, frozen):
    """freeze every free atom other than the centre farther than radius from
    it (minimum image); returns the new frozen mask"""
    pos = np.asarray(pos, dtype=np.float64)
    fr = np.array(frozen, dtype=np.int32, copy=True)
    d = W.min_image(pos - pos[:, centre_atom:centre_atom + 1], lat)
    far = (d * d).sum(0) > radius * radius
    far[centre_atom] = False
    fr[far & (fr == 0)] = 1
    return fr


def start_displacement(pos, frozen, uniform, d0, lat=None, centre=None, gauss_width=None):
    """makeRandomDisplacement, or makeGaussDisplacement when gauss_width is
    given: three uniforms per free atom in atom order, centre of mass of the
    displacement removed, the start direction normalised from the unweighted
    displacement.  Returns (displacement (3, n), unit start mode (3, n))."""
    pos = np.asarray(pos, dtype=np.float64)
    free = np.nonzero(np.asarray(frozen) == 0)[0]
    n = pos.shape[1]
    u = np.array([[uniform() - 0.5 for _ in range(3)] for _ in free]).reshape(-1, 3)
    w = np.ones(len(free))
    if gauss_width is not None:
        d = W.min_image(pos[:, free] - np.asarray(centre, dtype=np.float64)[:, None], lat)
        dist2 = (d * d).sum(0)
        w = np.exp(-dist2 / (2.0 * gauss_width ** 2))
    u -= u.mean(0, keepdims=True)         # centreOfMass over the 3 components separately
    mode = np.zeros((3, n))
    disp = np.zeros((3, n))
    norm = np.sqrt((u * u).sum())
    mode[:, free] = (u / norm).T
    disp[:, free] = (u * d0 * w[:, None]).T
    return disp, mode


class RegionProposer:
    """A SaddleCampaign proposer (callable(campaign, w) -> x0, mode, frozen,
    centre) with the reference's region styles and start displacements."""

    def __init__(self, style=2, region_radius=6.0, d0=0.1, gauss_width=None, species=None, art_types=None,
                 coord_distance=None, ideal_coord=None, seed=0):
        self.style, self.radius, self.d0, self.width = style, region_radius, d0, gauss_width
        self.species, self.art_types = species, art_types
        self.coord_distance, self.ideal_coord = coord_distance, ideal_coord
        self.rng = np.random.default_rng(seed)

    def __call__(self, camp, w):
        c = camp.cells[w]
        basin = camp.basins[w]
        frozen = np.zeros(c.n, dtype=np.int32)
        uni = self.rng.random
        centre = None
        if self.style != 1 or self.width is not None:
            j = choose_centre(self.style, basin, c.types, frozen, c.lat, uni, self.species, self.art_types,
                              self.coord_distance, self.ideal_coord)
            centre = basin[:, j].copy()
            if self.style != 1:
                frozen = sphere_region(basin, c.lat, j, self.radius, frozen)
        disp, mode = start_displacement(basin, frozen, uni, self.d0, c.lat, centre, self.width)
        return basin + disp, mode, frozen, centre if centre is not None else np.zeros(3)
