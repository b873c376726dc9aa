"""The allocation-free force entry point (akmc_field_forces_noalloc) against
akmc_field_forces on the same field and positions, through the C ABI: two
calls in a row, the second on a smaller cell, each bitwise equal to the
allocating entry point (the same kernels and variant run on both), with the
overflow flags clear; sizes past the stated maximum and non-Sutton-Chen
fields are refused before any launch.  Device buffers come from the HIP
runtime the library itself links (libamdhip64.so.7, through ctypes)."""
import ctypes as C

import numpy as np
import pytest

from akmc_amd import workloads as W

pytestmark = pytest.mark.gpu
CU = W.SC_CU


def _lib():
    from akmc_amd import lib as L
    lib = L.load()
    lib.akmc_field_forces_noalloc.argtypes = [C.c_void_p] + [C.c_void_p] * 4 + [C.c_int, C.POINTER(C.c_double)] + \
        [C.c_void_p] * 6
    lib.akmc_field_forces_noalloc.restype = C.c_int
    lib.akmc_noalloc_scratch_bytes.restype = C.c_ulong
    return L, lib


def _reference(L, f, c):
    """akmc_field_forces (host arrays; it wraps its copies of the positions),
    without stress, as the allocation-free entry point runs"""
    return f.calculate_forces(c.pos.copy(), c.types, c.lat, stress=False)


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the library's own HIP runtime"""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _noalloc(lib, fh, c):
    hip = _Hip()
    try:
        xs = [hip.put(np.ascontiguousarray(c.pos[k], dtype=np.float64)) for k in range(3)]
        typ = hip.put(np.ascontiguousarray(c.types, dtype=np.int32))
        nan = np.full(c.n, np.nan)
        fs = [hip.put(nan) for _ in range(3)]
        en = hip.put(np.full(6, np.nan))       # akmc_energy: 6 doubles
        flags = hip.put(np.full(1, -1, np.int32))
        lat = np.ascontiguousarray(c.lat, dtype=np.float64)
        rc = lib.akmc_field_forces_noalloc(fh, xs[0], xs[1], xs[2], typ, c.n, lat.ctypes.data_as(C.POINTER(C.c_double)),
                                           fs[0], fs[1], fs[2], en, flags, None)
        assert rc == 0, lib.akmc_last_error().decode()
        assert hip.h.hipDeviceSynchronize() == 0
        frc = np.stack([hip.get(p, (c.n,), np.float64) for p in fs])
        return frc, hip.get(en, (6,), np.float64), int(hip.get(flags, (1,), np.int32)[0])
    finally:
        hip.free()


def test_noalloc_forces_equal_allocating_entry_point():
    L, lib = _lib()
    print(f"scratch array: {lib.akmc_noalloc_scratch_bytes() / 2**20:.1f} MiB per GPU", flush=True)
    f = L.Field(L.sc_potentials_text([("Cu", "Cu", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"])], 2 * CU["a"]))
    big = W.sc_cell(6, 6, 6, n_vac=2, jitter=0.02, seed=11)    # 862 atoms
    small = W.plain_sc_workload(seed=1)                         # 255 atoms
    f.setup(big.lat, big.n)
    for name, c in (("862-atom cell", big), ("255-atom cell", small)):
        ref = _reference(L, f, c)
        frc, en, flags = _noalloc(lib, f._h, c)
        e = ref["energy"]
        print(f"{name}: flags {flags}; max|dF| vs akmc_field_forces {np.abs(frc - ref['force']).max():.3e}; "
              f"E {en[0]:.12f} vs {e['total']:.12f}", flush=True)
        assert flags == 0
        assert np.array_equal(frc, ref["force"])
        assert en[0] == e["total"] and en[3] == e["many"] and en[4] == e["vdw"]


def test_noalloc_refuses_what_it_cannot_serve():
    L, lib = _lib()
    f = L.Field(L.sc_potentials_text([("Cu", "Cu", CU["eps"], CU["a"], CU["n"], CU["m"], CU["c"])], 2 * CU["a"]))
    c = W.plain_sc_workload(seed=1)
    lat = np.ascontiguousarray(c.lat, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))
    dummy = C.c_void_p(16)  # never dereferenced: the calls below are refused before any launch
    args = [dummy] * 4
    # not set up
    rc = lib.akmc_field_forces_noalloc(f._h, *args, c.n, lat, dummy, dummy, dummy, None, None, None)
    assert rc == L.EINVAL and "akmc_field_setup" in lib.akmc_last_error().decode()
    f.setup(c.lat, c.n)
    rc = lib.akmc_field_forces_noalloc(f._h, *args, 8193, lat, dummy, dummy, dummy, None, None, None)
    assert rc == L.ECAPACITY and "8192" in lib.akmc_last_error().decode()
    ion = L.Field(L.buck_potentials_text(["M", "X"], [2.0, -2.0], [("M", "X") + W.ROCKSALT_BUCK[("M", "X")]], 5.0))
    rc = lib.akmc_field_forces_noalloc(ion._h, *args, 10, lat, dummy, dummy, dummy, None, None, None)
    assert rc == L.EUNSUPPORTED
