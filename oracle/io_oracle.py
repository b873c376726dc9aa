"""CPU restatement of the reference's text formats and KMC event packet.

TEST INFRASTRUCTURE ONLY: the checker of the library's akmc_io_* and
akmc_kmc_event_* functions.  Nothing in the product path imports it.

C++ iostream output with showpoint|fixed and precision p is printf's %.pf;
setw(w) right-aligns in w characters; a double streamed with no flags set
uses %g with 6 significant digits.  Restated here with Python's %-format:
Basis::dumpBasis (Common/Basis.cpp:600-649), Basis::printVectors
(:1091-1112), Basis::inputBasis (:356-501), runKMC's restart file
(kmc/KineticMonteCarlo.cpp:210-218), writeEventList (:1561-1616) with
Event::writeCentre / writeSaddle / writeBasin2 (kmc/Event.cpp:216-292),
and the event packet of sendEvent (:1376-1461) / Event(int*, float*, ...)
(kmc/Event.cpp:66-105).  Pinned by tests/test_io_formats.py: the
reference's own dumpBasis output docs/examples/gdc/basis (kept as
tests/golden/gdc_basis.txt) reproduced byte for byte, and hand-worked
event files and packets.
"""
from __future__ import annotations

import numpy as np


def _lattice(lat):
    return "".join("%15.10f%15.10f%15.10f\n" % tuple(lat[3 * r:3 * r + 3]) for r in range(3))


def dump_basis(title, names, labels, pos, lat, time, energy, iteration, header=True):
    n = pos.shape[1]
    out = []
    if header:
        out.append(f"{title}\n0    3   {n}\n")
    out.append("%15s%d     %d    0      3   %.10f     %.10f\n" % ("timeStep ", iteration, n, time, energy))
    out.append(_lattice(lat))
    for i in range(n):
        out.append(f"{names[labels[i]]}      {i + 1}\n")
        out.append("  %.10f  %.10f  %.10f\n" % (pos[0, i], pos[1, i], pos[2, i]))
    return "".join(out)


def restart(title, names, labels, pos, lat, time, energy, iteration, header=False):
    n = pos.shape[1]
    return f"restart {title}\n 0   3 {n} {n}\n" + dump_basis(title, names, labels, pos, lat, time, energy,
                                                              iteration, header)


def read_basis(text, names):
    lines = text.split("\n")
    title = lines[0]
    w = lines[1].split()
    levcfg, n = int(w[0]), int(w[2])
    at = 2
    it, t = 0, 0.0
    w = lines[at].split()
    if w[0] == "timeStep":
        it, t = int(w[1]), float(w[5])
        at += 1
    lat = [float(v) for r in range(3) for v in lines[at + r].split()[:3]]
    at += 3
    labels, pos = [], np.zeros((3, n))
    for j in range(n):
        labels.append(names.index(lines[at].split()[0]))
        pos[:, j] = [float(v) for v in lines[at + 1].split()[:3]]
        at += 2 + (1 if levcfg >= 1 else 0) + (1 if levcfg >= 2 else 0)
    return dict(title=title, levcfg=levcfg, labels=labels, pos=pos, lat=lat, iteration=it, time=t)


def _g6(x):
    return "%g" % x


def events_file(names, labels, basin1, lat, basin1_energy, events, chosen, time, fp32=True):
    """events: list of dicts with act, centre (3,), saddle_energy, saddle
    (3, n), basin2_energy, basin2 (3, n)"""
    f = (lambda v: float(np.float32(v))) if fp32 else float
    n = basin1.shape[1]
    out = [f"events   {len(events)} chosen event {chosen} time {_g6(time)}\n",
           "%15s%.10f\n" % ("basin1 ", basin1_energy), _lattice(lat)]
    for i in range(n):
        out.append("%5s    %d\n%15.10f%15.10f%15.10f\n" % (names[labels[i]], i, *basin1[:, i]))
    for e, ev in enumerate(events):
        out.append(f" event {e}\n")
        out.append(" activation energy %.10f\n" % f(ev["act"]))
        out.append("%15s%.7f %.7f %.7f\n" % ("centre ", *[f(c) for c in ev["centre"]]))
        for tag, ek, pk in (("saddle ", "saddle_energy", "saddle"), ("basin2 ", "basin2_energy", "basin2")):
            out.append("%15s%.10f  %d\n" % (tag, f(ev[ek]), n))
            P = ev[pk]
            for i in range(n):
                out.append("%5s    %d\n%15.10f%15.10f%15.10f\n" % (names[labels[i]], i, f(P[0, i]), f(P[1, i]),
                                                                   f(P[2, i])))
    return "".join(out)


def event_pack(global_no, basin2, saddle, e1, e2, es, ea, centre):
    n = len(global_no)
    vecf = np.zeros(6 * n + 7, dtype=np.float32)
    for i in range(n):
        vecf[6 * i:6 * i + 3] = basin2[:, i]
        vecf[6 * i + 3:6 * i + 6] = saddle[:, i]
    vecf[6 * n:6 * n + 7] = [e1, e2, es, ea, centre[0], centre[1], centre[2]]
    return np.asarray(global_no, dtype=np.int32), vecf
