"""Writes tests/golden/c2_art_oracle.json: one ART search of the rock-salt
Ewald cell (510 ions, charges +-2, Buckingham, eta = 0.35 1/A, rc = 10 A
over explicit images, |n| < 8) run by the CPU FP64 oracle, for the GPU test
that starts the device ART from the same basin and start.

The basin is the oracle's FIRE relax of the seeded cell; the start displaces
the 12 cations nearest the cation vacancy by a seeded 0.1 A random
displacement (atoms beyond 6 A frozen).  The oracle's ART takes a few
minutes of serial FP64 force calls, which is why its result is kept as a
fixture instead of being recomputed by the GPU test.

    python tests/golden/make_c2_art_fixture.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from akmc_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

SEED_CELL, SEED_DISP = 5, 100
ART = dict(initTimeStep=0.0005, maxTimeStep=0.002, forceTol=1e-2, parDamp=0.25, numLanczosVectors=25)


def c2_setup():
    c = W.rocksalt_cell(seed=SEED_CELL)
    pots = [("M", "X") + W.ROCKSALT_BUCK[("M", "X")], ("X", "X") + W.ROCKSALT_BUCK[("X", "X")],
            ("M", "M") + W.ROCKSALT_BUCK[("M", "M")]]
    return c, pots


def main():
    c, pots = c2_setup()
    ew = O.Ewald(c.lat, 0.35, gcell=8)
    of = O.ion_field(["M", "X"], pots, rc=10.0, ewald=ew, min_image=False, lat=c.lat)
    h = O.FieldHandle(of, 1, c.types, c.lat, charges=c.charges)
    r = O.fire_relax(h, c.pos)
    assert r["status"] == 0
    basin = r["pos"]
    site = c.vacancies[0]
    d = W.min_image(basin - site[:, None], c.lat)
    rr = np.sqrt((d * d).sum(0))
    frozen = (rr >= 6.0).astype(np.int32)
    cations = np.nonzero(c.types == 0)[0]
    movers = cations[np.argsort(rr[cations])[:12]]
    disp, mode = W.local_displacement(W.Cell(basin, c.types, c.lat), movers, 0.1, SEED_DISP)
    cn = O.art_defaults()
    for k, v in ART.items():
        setattr(cn, k, v)
    m = mode[:, frozen == 0].T.reshape(-1)
    a = O.art_search(h, basin + disp, frozen, m / np.linalg.norm(m), cn)
    out = dict(cell_seed=SEED_CELL, disp_seed=SEED_DISP, art=ART, basin=basin.tolist(),
               basin_energy=r["energy"], art_status=a["status"], art_iters=a["iters"],
               saddle_energy=a["energy"], barrier=a["energy"] - r["energy"], eigval=a["eigval"],
               saddle_free=a["pos"][:, frozen == 0].tolist(), force_calls=int(h.ncalls))
    with open(os.path.join(HERE, "c2_art_oracle.json"), "w") as fh:
        json.dump(out, fh)
    print(f"basin {r['energy']:.8f} eV ({r['iters']} FIRE iterations); ART status {a['status']} "
          f"barrier {out['barrier']:.6f} eV eig {a['eigval']:.5f} iters {a['iters']} calls {h.ncalls}")


if __name__ == "__main__":
    main()
