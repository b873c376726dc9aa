"""The C-ABI library without a GPU: it loads, exports every symbol that
include/*.h declares, and its host-side potentials parser follows the
reference's readPotential rules (no device calls are made here)."""
import ctypes as C
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(akmc_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_declared_symbols():
    from akmc_amd import lib
    lib.load()
    out = subprocess.run(["nm", "-D", "--defined-only", lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    declared = _declared()
    assert len(declared) >= 20
    missing = declared - exported
    assert not missing, missing


def _create(text):
    from akmc_amd import lib
    L = lib.load()
    h = C.c_void_p()
    rc = L.akmc_field_create(text.encode(), C.byref(h))
    return rc, h, L.akmc_last_error().decode()


def test_parser_accepts_reference_example():
    text = open(os.path.join(ROOT, "tests", "golden", "gdc_potentials.txt")).read()
    rc, h, err = _create(text)
    assert rc == 0, err
    from akmc_amd import lib
    assert lib.load().akmc_field_kind(h) == 1
    lib.load().akmc_field_destroy(h)


def test_parser_sutton_chen_and_errors():
    from akmc_amd import lib
    text = lib.sc_potentials_text([("Cu", "Cu", 0.012382, 3.61, 9.0, 6.0, 39.432)], 7.22)
    rc, h, err = _create(text)
    assert rc == 0 and lib.load().akmc_field_kind(h) == 0
    lib.load().akmc_field_destroy(h)
    rc, h, err = _create("cutoff 5\nfoo 1\nclose\n")
    assert rc == lib.EINVAL and "unrecognised key word" in err
    rc, h, err = _create("cutoff 5\nspecies 1\nCu 1 0\nmanybody 1 ev\ngupt\nCu Cu 1 2 3 4 5\nclose\n")
    assert rc == lib.EUNSUPPORTED
    rc, h, err = _create("cutoff 5\nclose\n")
    assert rc == lib.EINVAL


def test_python_binding_fails_loudly_without_library(monkeypatch, tmp_path):
    from akmc_amd import lib
    monkeypatch.setattr(lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(lib, "_lib", None)
    with pytest.raises(ImportError):
        lib.load()


def test_status_array_layout():
    """The numpy view of akmc_walker_status reads every field at its C offset."""
    import ctypes as C
    import numpy as np
    from akmc_amd import lib as L
    st = (L.WalkerStatus * 3)()
    st[1].done, st[1].energy, st[2].lambda2, st[2].force_calls, st[0].verified = 1, -3.5, 1.25, 77, 1
    a = np.frombuffer(st, dtype=L._STATUS_DTYPE)
    assert L._STATUS_DTYPE.itemsize == C.sizeof(L.WalkerStatus)
    assert list(a["done"]) == [0, 1, 0] and list(a["force_calls"]) == [0, 0, 77]
    assert a["energy"][1] == -3.5 and a["lambda2"][2] == 1.25 and a["verified"][0] == 1
