"""The reference's text formats and KMC event packet on the host (no
device calls): the library's writers and readers (akmc_io_*,
akmc_kmc_event_*) and their oracle (oracle/io_oracle.py), pinned by the
reference's own dumpBasis output docs/examples/gdc/basis (kept byte for
byte as tests/golden/gdc_basis.txt) and by hand-worked files and packets."""
import math
import os

import numpy as np
import pytest

from oracle import io_oracle as IO
from oracle import kmc_oracle as K

HERE = os.path.dirname(os.path.abspath(__file__))
GDC = os.path.join(HERE, "golden", "gdc_basis.txt")
GDC_SPECIES = ["Ce", "Gd", "O"]


@pytest.fixture(scope="module")
def lib():
    from akmc_amd import lib as L
    L.load()
    return L


def _gdc_energy():
    return float(open(GDC).read().split("\n")[2].split()[6])


def test_oracle_reproduces_reference_basis_file():
    text = open(GDC).read()
    b = IO.read_basis(text, GDC_SPECIES)
    assert b["levcfg"] == 0 and b["iteration"] == 1 and b["time"] == 0.0 and len(b["labels"]) == 2549
    out = IO.dump_basis(b["title"], GDC_SPECIES, b["labels"], b["pos"], b["lat"], b["time"], _gdc_energy(),
                        b["iteration"])
    assert out == text


def test_library_basis_round_trip_is_byte_exact(lib, tmp_path):
    b = lib.read_basis(GDC, GDC_SPECIES)
    assert b["title"] == open(GDC).read().split("\n")[0]
    assert b["pos"].shape == (3, 2549) and b["pos"][0, 0] == 1.4168603868 and b["lat"][4] == 32.7232154959
    assert (b["iteration"], b["time"], b["levcfg"]) == (1, 0.0, 0)
    out = tmp_path / "basis"
    lib.dump_basis(out, b["title"], GDC_SPECIES, b["labels"], b["pos"], b["lat"], b["time"], _gdc_energy(),
                   b["iteration"])
    assert out.read_bytes() == open(GDC, "rb").read()


def test_archive_append_and_restart(lib, tmp_path):
    rng = np.random.default_rng(1)
    names = ["Cu", "Ag"]
    lab = rng.integers(0, 2, 7)
    pos = rng.normal(0, 3, (3, 7))
    lat = np.diag([9.5, 10.25, 11.0]).reshape(9)
    arc = tmp_path / "archive"
    lib.dump_basis(arc, "run", names, lab, pos, lat, 0.0, -12.5, 1, header=True)
    lib.dump_basis(arc, "run", names, lab, pos + 0.1, lat, 3.5e-7, -12.75, 2, header=False, append=True)
    exp = IO.dump_basis("run", names, lab, pos, lat, 0.0, -12.5, 1) + \
        IO.dump_basis("run", names, lab, pos + 0.1, lat, 3.5e-7, -12.75, 2, header=False)
    assert arc.read_text() == exp
    rs = tmp_path / "restart"
    lib.write_restart(rs, "run", names, lab, pos, lat, 1.0e-6, -12.5, 3)
    assert rs.read_text() == IO.restart("run", names, lab, pos, lat, 1.0e-6, -12.5, 3)
    # the restart file reads back (title "restart run", iteration, time, positions to 1e-10)
    back = lib.read_basis(rs, names)
    assert back["title"] == "restart run" and back["iteration"] == 3 and back["time"] == 1.0e-6
    assert np.abs(back["pos"] - pos).max() <= 5e-11 and list(back["labels"]) == list(lab)
    with pytest.raises(lib.AkmcError, match="no atom of type"):
        lib.read_basis(rs, ["Cu"])


EXPECTED_EVENTS = (
    "events   1 chosen event 0 time 1.5e-10\n"
    "        basin1 -3.2500000000\n"
    "  10.0000000000   0.0000000000   0.0000000000\n"
    "   0.0000000000  10.0000000000   0.0000000000\n"
    "   0.0000000000   0.0000000000  10.0000000000\n"
    "   Cu    0\n   0.0000000000   0.0000000000   0.0000000000\n"
    "   Cu    1\n   1.5000000000   0.0000000000   0.0000000000\n"
    " event 0\n"
    " activation energy 0.5000000000\n"
    "        centre 1.0000000 2.0000000 3.0000000\n"
    "        saddle -2.7500000000  2\n"
    "   Cu    0\n   0.2500000000   0.0000000000   0.0000000000\n"
    "   Cu    1\n   1.5000000000   0.0000000000   0.0000000000\n"
    "        basin2 -3.0000000000  2\n"
    "   Cu    0\n   0.5000000000   0.0000000000   0.0000000000\n"
    "   Cu    1\n   1.5000000000   0.0000000000   0.0000000000\n")


def _two_atom_event():
    b1 = np.array([[0.0, 1.5], [0.0, 0.0], [0.0, 0.0]])
    ev = dict(act=0.5, centre=[1.0, 2.0, 3.0], saddle_energy=-2.75, basin2_energy=-3.0,
              saddle=np.array([[0.25, 1.5], [0.0, 0.0], [0.0, 0.0]]),
              basin2=np.array([[0.5, 1.5], [0.0, 0.0], [0.0, 0.0]]))
    return b1, ev


def test_events_file_known_answer(lib, tmp_path):
    b1, ev = _two_atom_event()
    lat = np.diag([10.0] * 3).reshape(9)
    assert IO.events_file(["Cu"], [0, 0], b1, lat, -3.25, [ev], 0, 1.5e-10) == EXPECTED_EVENTS
    p = tmp_path / "events.1"
    lib.write_events(p, ["Cu"], [0, 0], b1, lat, -3.25, [ev], 0, 1.5e-10)
    assert p.read_text() == EXPECTED_EVENTS


@pytest.mark.parametrize("fp32", [True, False])
def test_events_file_matches_oracle(lib, tmp_path, fp32):
    rng = np.random.default_rng(5)
    n, names = 9, ["Cu", "Ni", "Ag"]
    lab = rng.integers(0, 3, n)
    b1 = rng.normal(0, 4, (3, n))
    lat = np.diag([12.5, 13.0, 14.75]).reshape(9)
    evs = [dict(act=float(rng.uniform(0.1, 2)), centre=rng.normal(0, 3, 3), saddle_energy=float(rng.normal(-50, 1)),
                basin2_energy=float(rng.normal(-51, 1)), saddle=b1 + rng.normal(0, 0.3, (3, n)),
                basin2=b1 + rng.normal(0, 0.6, (3, n))) for _ in range(3)]
    p = tmp_path / "events.7"
    lib.write_events(p, names, lab, b1, lat, -51.234567, evs, 2, 4.56e-3, fp32=fp32)
    assert p.read_text() == IO.events_file(names, lab, b1, lat, -51.234567, evs, 2, 4.56e-3, fp32=fp32)


def test_event_packet(lib):
    rng = np.random.default_rng(3)
    n = 6
    g = rng.permutation(40)[:n]
    b2, sd = rng.normal(0, 5, (3, n)), rng.normal(0, 5, (3, n))
    e1, e2, es, ea, cen = -100.123456789, -100.3, -99.6, 0.523456789, [1.5, -2.25, 3.125]
    vi, vf = lib.event_pack(g, b2, sd, e1, e2, es, ea, cen)
    ei, ef = IO.event_pack(g, b2, sd, e1, e2, es, ea, cen)
    assert np.array_equal(vi, ei) and np.array_equal(vf, ef)
    # layout by hand: atom 2's saddle y, the activation energy, the centre z
    assert vf[6 * 2 + 4] == np.float32(sd[1, 2]) and vf[6 * n + 3] == np.float32(ea) and vf[6 * n + 6] == 3.125
    u = lib.event_unpack(vi, vf, 1.0e13, 1000.0)
    assert np.array_equal(u["global_no"], g)
    assert np.array_equal(u["basin2"], b2.astype(np.float32).astype(np.float64))
    assert u["activation_energy"] == float(np.float32(ea)) != ea      # the packet's FP32 rounding
    assert u["rate"] == K.rate(float(np.float32(ea)), 1.0e13, 1000.0)
    assert math.isclose(u["rate"], 1.0e13 * math.exp(-ea / (8.617333262e-5 * 1000.0)), rel_tol=1e-6)


def test_cpp_basis_round_trip(tmp_path):
    """The C++ host classes: Basis::inputBasis then Basis::dumpBasis
    (include/akmc.hpp, through tests/cpp/host_driver.cpp) reproduce the
    reference's file byte for byte."""
    import subprocess
    drv = os.path.join(os.path.dirname(HERE), "akmc_amd", "_build", "host_driver")
    out = tmp_path / "basis"
    r = subprocess.run([drv, "basis", " ".join(GDC_SPECIES), GDC, str(out)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == open(GDC, "rb").read()
