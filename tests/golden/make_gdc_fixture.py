"""Writes tests/golden/gdc_example.json from the reference's example input
data (docs/examples/gdc/{basis,potentials}): cell, species, charges,
Buckingham parameters, positions and the total energy the reference printed
in the basis header (docs/examples/gdc/basis:3).  Also keeps the basis file
itself, byte for byte, as tests/golden/gdc_basis.txt: it is a
Basis::dumpBasis output (its "timeStep" header line), the golden file of the
basis reader and writer.  Run in the build container, where the reference
tree is mounted; the JSON and the copy are the fixtures.
"""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/docs/examples/gdc"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gdc_example.json")
BASIS_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gdc_basis.txt")


def main():
    pot = [l.split() for l in open(os.path.join(REF, "potentials")) if l.strip()]
    cutoff = float(pot[0][1])
    nspec = int(pot[1][1])
    species, charges_by = [], {}
    for k in range(nspec):
        w = pot[2 + k]
        species.append(w[0])
        charges_by[w[0]] = float(w[2])
    buck = []
    i = 2 + nspec + 1
    while pot[i][0].lower() != "close":
        if pot[i][0].lower().startswith("buck"):
            w = pot[i + 1]
            buck.append([w[0], w[1], float(w[2]), float(w[3]), float(w[4])])
            i += 2
        else:
            i += 1
    lines = open(os.path.join(REF, "basis")).read().split("\n")
    n = int(lines[1].split()[2])
    hdr = lines[2].split()
    energy = float(hdr[6])
    lat = [float(v) for l in lines[3:6] for v in l.split()[:3]]
    xs, ys, zs, types, q = [], [], [], [], []
    for a in range(n):
        s = lines[6 + 2 * a].split()[0]
        x, y, z = [float(v) for v in lines[7 + 2 * a].split()[:3]]
        xs.append(x), ys.append(y), zs.append(z)
        types.append(species.index(s))
        q.append(charges_by[s])
    out = dict(source="docs/examples/gdc (basis header energy: basis:3)", cutoff=cutoff,
               ewald_precis=1.0e-6, species=species, buckingham=buck, lat=lat, types=types, charges=q,
               pos=[xs, ys, zs], energy=energy)
    json.dump(out, open(OUT, "w"))
    print("wrote", OUT, n, "ions, energy", energy)
    with open(os.path.join(REF, "basis"), "rb") as src, open(BASIS_OUT, "wb") as dst:
        dst.write(src.read())
    print("wrote", BASIS_OUT)


if __name__ == "__main__":
    main()
