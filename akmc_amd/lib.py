"""ctypes binding of libakmc_hip.so (include/akmc.h).

This is the only way Python reaches the device: there is no CPU fallback.
If the HIP library is missing the import of any binding raises.
The classes mirror the reference's Field plugin API
(potentials/metal/Field.h:137-217, potentials/pair/Field.h:157-234):
``Field.read_potential`` (readPotential), ``setup``, ``calculate_forces``
(calculateForces, positions wrapped in place) and ``calculate_energy``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libakmc_hip.so")

SUCCESS, FAILED, INTERRUPTED, TERMINATED = 0, 1, 3, 4
EINVAL, EHIP, ECAPACITY, EUNSUPPORTED = -1, -2, -3, -4

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


class Energy(C.Structure):
    """akmc_energy: the Energy record (potentials/metal/Energy.h:17-37)."""
    _fields_ = [("total", C.c_double), ("rcp", C.c_double), ("real", C.c_double), ("many", C.c_double),
                ("vdw", C.c_double), ("ext", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RelaxParams(C.Structure):
    """akmc_relax_params: RelaxCntrl field names (Common/RelaxCntrl.h:15-48)."""
    _fields_ = [("maxIter", C.c_int), ("normTol", C.c_double), ("forceTol", C.c_double), ("N_min", C.c_int),
                ("initTimeStep", C.c_double), ("alphaDecrease", C.c_double), ("stepIncrease", C.c_double),
                ("stepDecrease", C.c_double), ("alphaStart", C.c_double), ("maxTimeStep", C.c_double),
                ("timeStepUnit", C.c_double), ("method", C.c_int), ("verletIntegration", C.c_int),
                ("N_delaySteps", C.c_int), ("N_maxUphill", C.c_int), ("initialDelay", C.c_int),
                ("minTimeStep", C.c_double)]


# Relax::RelaxStructure's method strings (optim/Relax.cpp:25-40); any other
# string is a single force and energy evaluation
RELAX_FIRE, RELAX_FIRE2, RELAX_NONE = 0, 1, 2


def relax_method(name: str) -> int:
    return {"fire": RELAX_FIRE, "fire2": RELAX_FIRE2}.get(name, RELAX_NONE)


class DimerParams(C.Structure):
    """akmc_dimer_params: DimerCntrl field names (Common/DimerCntrl.h:18-88)."""
    _fields_ = [("deltaTheta", C.c_double), ("dimerDisplacement", C.c_double), ("maxDeltaX", C.c_double),
                ("maxRotations", C.c_int), ("maxParIter", C.c_int), ("maxEigenvalue", C.c_double),
                ("normTol", C.c_double), ("forceTol", C.c_double), ("N_min", C.c_int),
                ("alphaDecrease", C.c_double), ("stepIncrease", C.c_double), ("stepDecrease", C.c_double),
                ("alphaStart", C.c_double), ("initTimeStep", C.c_double), ("maxTimeStep", C.c_double),
                ("maxStep", C.c_double), ("timeStepUnit", C.c_double)]


class ARTParams(C.Structure):
    """akmc_art_params: ARTCntrl field names (Common/ArtCntrl.h:14-74)."""
    _fields_ = [("numLanczosVectors", C.c_int), ("deltaX", C.c_double), ("maxEigenvalue", C.c_double),
                ("eigValTol", C.c_double), ("parDamp", C.c_double), ("maxParIter", C.c_int),
                ("normTol", C.c_double), ("forceTol", C.c_double), ("N_min", C.c_int),
                ("initTimeStep", C.c_double), ("alphaDecrease", C.c_double), ("stepIncrease", C.c_double),
                ("stepDecrease", C.c_double), ("alphaStart", C.c_double), ("maxTimeStep", C.c_double),
                ("timeStepUnit", C.c_double)]


class WalkerStatus(C.Structure):
    _fields_ = [("kind", C.c_int), ("status", C.c_int), ("done", C.c_int), ("iterations", C.c_int),
                ("force_calls", C.c_int), ("energy", C.c_double), ("curvature", C.c_double),
                ("verified", C.c_int), ("energy32", C.c_double), ("curvature32", C.c_double),
                ("energy64", C.c_double), ("fmax64", C.c_double), ("lambda1", C.c_double),
                ("lambda2", C.c_double), ("new_basin", C.c_int), ("basin2_status", C.c_int),
                ("basin2_force_calls", C.c_int), ("basin2_iterations", C.c_int),
                ("energy_basin2", C.c_double), ("energy_basin2_64", C.c_double),
                ("basin2_max_disp", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# numpy view of WalkerStatus with the C layout's offsets
_STATUS_DTYPE = np.dtype({"names": [k for k, _ in WalkerStatus._fields_],
                          "formats": [np.int32 if t is C.c_int else np.float64 for _, t in WalkerStatus._fields_],
                          "offsets": [getattr(WalkerStatus, k).offset for k, _ in WalkerStatus._fields_],
                          "itemsize": C.sizeof(WalkerStatus)})


class Basin2Params(C.Structure):
    """akmc_basin2_params: the basin-2 keys of the reference's kmc{} block
    (basin2Delta, kmcBasinRadius, kmcMaxRadius; kmc/KineticMonteCarlo.cpp:1672,1690)."""
    _fields_ = [("basin2Delta", C.c_double), ("kmcBasinRadius", C.c_double), ("kmcMaxRadius", C.c_double)]


def basin2_defaults(**kw) -> Basin2Params:
    p = Basin2Params()
    load().akmc_basin2_defaults(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def relax_defaults(**kw) -> RelaxParams:
    """RelaxCntrl defaults; method may be given as the reference's string"""
    p = RelaxParams()
    load().akmc_relax_defaults(C.byref(p))
    for k, v in kw.items():
        if k == "method" and isinstance(v, str):
            v = relax_method(v)
        setattr(p, k, v)
    return p


def art_defaults(**kw) -> ARTParams:
    p = ARTParams()
    load().akmc_art_defaults(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def dimer_defaults(**kw) -> DimerParams:
    p = DimerParams()
    load().akmc_dimer_defaults(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class AkmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"akmc error {code}: {msg}")
        self.code = code


_lib = None


def load() -> C.CDLL:
    """Load libakmc_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m akmc_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.akmc_last_error.restype = C.c_char_p
    L.akmc_field_create.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
    L.akmc_field_destroy.argtypes = [C.c_void_p]
    L.akmc_field_kind.argtypes = [C.c_void_p]
    L.akmc_field_set_ewald.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int]
    L.akmc_field_set_skin.argtypes = [C.c_void_p, C.c_double]
    L.akmc_field_set_list_density.argtypes = [C.c_void_p, C.c_double]
    L.akmc_field_set_precision.argtypes = [C.c_void_p, C.c_int]
    L.akmc_field_setup.argtypes = [C.c_void_p, _D, C.c_int]
    L.akmc_field_forces.argtypes = [C.c_void_p, _D, _D, _D, _D, _D, _D, _D, _D, _D, _I, _I, C.c_int,
                                    C.POINTER(Energy)]
    L.akmc_field_energy.argtypes = [C.c_void_p, _D, _D, _D, _D, _D, _I, C.c_int, C.POINTER(Energy)]
    L.akmc_field_num_kvectors.argtypes = [C.c_void_p]
    L.akmc_batch_create.argtypes = [C.c_void_p, C.c_int, _I, _D, C.POINTER(C.c_void_p)]
    L.akmc_batch_destroy.argtypes = [C.c_void_p]
    L.akmc_batch_upload.argtypes = [C.c_void_p, _D, _D, _D, _I, _D, _I]
    L.akmc_batch_download.argtypes = [C.c_void_p, _D, _D, _D, _D, _D, _D]
    L.akmc_batch_forces.argtypes = [C.c_void_p, C.POINTER(Energy)]
    L.akmc_batch_forces_repeat.argtypes = [C.c_void_p, C.c_int, _D]
    L.akmc_batch_total_atoms.argtypes = [C.c_void_p]
    L.akmc_batch_rebuilds.argtypes = [C.c_void_p]
    L.akmc_batch_rebuilds.restype = C.c_longlong
    L.akmc_relax_defaults.argtypes = [C.POINTER(RelaxParams)]
    L.akmc_dimer_defaults.argtypes = [C.POINTER(DimerParams)]
    L.akmc_batch_set_frozen.argtypes = [C.c_void_p, _I]
    L.akmc_batch_start_relax.argtypes = [C.c_void_p, _I, C.POINTER(RelaxParams)]
    L.akmc_batch_start_dimer.argtypes = [C.c_void_p, _I, _D, _D, _I, C.POINTER(DimerParams)]
    L.akmc_batch_run.argtypes = [C.c_void_p, C.c_int, C.c_int, _I]
    L.akmc_batch_run_async.argtypes = [C.c_void_p, C.c_int]
    L.akmc_batch_run_wait.argtypes = [C.c_void_p, _I, _I]
    L.akmc_batch_run_next.argtypes = [C.c_void_p, C.c_int]
    L.akmc_batch_set_harvest.argtypes = [C.c_void_p, C.c_int]
    L.akmc_batch_get_harvested_positions.argtypes = [C.c_void_p, C.c_int, _I, _D]
    L.akmc_batch_status.argtypes = [C.c_void_p, C.POINTER(WalkerStatus)]
    L.akmc_batch_get_positions.argtypes = [C.c_void_p, _D, _D, _D]
    L.akmc_batch_get_mode.argtypes = [C.c_void_p, _D]
    L.akmc_batch_time_phases.argtypes = [C.c_void_p, C.c_int, _D]
    L.akmc_batch_set_variant.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.akmc_batch_set_graphs.argtypes = [C.c_void_p, C.c_int]
    L.akmc_batch_set_search_zone.argtypes = [C.c_void_p, C.c_int]
    L.akmc_batch_set_logic_block.argtypes = [C.c_void_p, C.c_int]
    L.akmc_batch_time_search_phases.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
    L.akmc_batch_get_zone.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 4
    L.akmc_batch_set_search_variant.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.akmc_batch_autotune_search.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                             C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.akmc_batch_get_walkers_positions.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), _D]
    L.akmc_batch_kernel_timer.argtypes = [C.c_void_p, C.c_int]
    L.akmc_batch_kernel_time.argtypes = [C.c_void_p, _D, C.POINTER(C.c_longlong)]
    L.akmc_batch_eval_counts.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
    L.akmc_batch_autotune.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), _D]
    L.akmc_batch_set_staging.argtypes = [C.c_void_p, C.c_int]
    L.akmc_batch_staging.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.akmc_batch_save_basin.argtypes = [C.c_void_p]
    L.akmc_art_defaults.argtypes = [C.POINTER(ARTParams)]
    L.akmc_batch_start_art.argtypes = [C.c_void_p, _I, _D, _D, _I, C.POINTER(ARTParams)]
    L.akmc_batch_set_verify.argtypes = [C.c_void_p, C.c_int, C.c_double]
    L.akmc_batch_search_energies.argtypes = [C.c_void_p, C.c_int, _D]
    L.akmc_batch_get_walker_positions.argtypes = [C.c_void_p, C.c_int, _D, _D, _D]
    L.akmc_batch_restore_basin.argtypes = [C.c_void_p, _I]
    L.akmc_basin2_defaults.argtypes = [C.POINTER(Basin2Params)]
    L.akmc_batch_set_basin2.argtypes = [C.c_void_p, C.POINTER(Basin2Params)]
    L.akmc_batch_get_walkers_saddles.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), _D]
    L.akmc_batch_set_basins.argtypes = [C.c_void_p, _I, _D]
    L.akmc_batch_start_relax_at.argtypes = [C.c_void_p, _I, _D, _I, C.POINTER(RelaxParams)]
    L.akmc_batch_cancel.argtypes = [C.c_void_p, _I]
    L.akmc_kmc_rate.argtypes = [C.c_double, C.c_double, C.c_double]
    L.akmc_kmc_rate.restype = C.c_double
    L.akmc_kmc_max_cap.argtypes = [C.c_double, C.c_double, C.c_double]
    L.akmc_kmc_max_cap.restype = C.c_double
    L.akmc_kmc_select.argtypes = [C.c_int, _D, C.c_double, _D]
    L.akmc_kmc_time_increase.argtypes = [C.c_double, C.c_double]
    L.akmc_kmc_time_increase.restype = C.c_double
    L.akmc_kmc_check_basin.argtypes = [C.c_int, _D, _D, _D, C.c_double, C.c_double]
    L.akmc_kmc_displaced_atoms.argtypes = [C.c_int, _D, _D, _D, C.c_double, _I]
    L.akmc_kmc_recycle.argtypes = [C.c_int, _D, C.c_int, _D, C.c_double, C.c_int, C.c_int, C.c_double, _D, _I,
                                   _I]
    L.akmc_kmc_update_excluded.argtypes = [C.c_int, _I, C.c_int, _I, C.c_int]
    _F = C.POINTER(C.c_float)
    L.akmc_kmc_event_pack.argtypes = [C.c_int, _I, _D, _D, C.c_double, C.c_double, C.c_double, C.c_double, _D,
                                      _I, _F]
    L.akmc_kmc_event_unpack.argtypes = [C.c_int, _I, _F, C.c_double, C.c_double, _I, _D, _D, _D, _D, _D]
    L.akmc_io_dump_basis.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_int, _I, _D, _D,
                                     C.c_double, C.c_double, C.c_int]
    L.akmc_io_write_restart.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_char_p, C.c_int, _I, _D, _D,
                                        C.c_double, C.c_double, C.c_int]
    L.akmc_io_basis_size.argtypes = [C.c_char_p, _I, _I]
    L.akmc_io_read_basis.argtypes = [C.c_char_p, C.c_char_p, C.c_int, _I, _D, _D, _I, _D, C.c_char_p, C.c_int]
    L.akmc_io_write_events.argtypes = [C.c_char_p, C.c_char_p, C.c_int, _I, _D, _D, C.c_double, C.c_int, _D, _D,
                                       _D, _D, _D, _D, C.c_int, C.c_double, C.c_int]
    _lib = L
    return L


def check(rc: int) -> int:
    if rc < 0:
        raise AkmcError(rc, load().akmc_last_error().decode())
    return rc


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(_I) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Field:
    """Force-field handle (the reference's compile-time Field plugin)."""

    def __init__(self, potentials_text: str):
        L = load()
        h = C.c_void_p()
        check(L.akmc_field_create(potentials_text.encode(), C.byref(h)))
        self._h = h
        self.kind = L.akmc_field_kind(h)

    @classmethod
    def read_potential(cls, path: str) -> "Field":
        with open(path) as fh:
            return cls(fh.read())

    def close(self):
        if getattr(self, "_h", None):
            load().akmc_field_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_ewald(self, eta, gcell):
        g = np.broadcast_to(np.asarray(gcell, dtype=int), (3,))
        check(load().akmc_field_set_ewald(self._h, float(eta), int(g[0]), int(g[1]), int(g[2])))

    def set_skin(self, skin):
        check(load().akmc_field_set_skin(self._h, float(skin)))

    def set_list_density(self, atoms_per_A3):
        """size later batches' neighbour lists for this number density"""
        check(load().akmc_field_set_list_density(self._h, float(atoms_per_A3)))

    def set_precision(self, bits):
        check(load().akmc_field_set_precision(self._h, int(bits)))

    def setup(self, lat, max_atoms):
        check(load().akmc_field_setup(self._h, _dp(_f64(lat)), int(max_atoms)))

    @property
    def num_kvectors(self):
        return load().akmc_field_num_kvectors(self._h)

    def calculate_forces(self, pos, types, lat, charges=None, stress=True):
        """Returns dict(pos=wrapped (3,N), force (3,N), stress (6,), energy dict)."""
        pos = np.array(pos, dtype=np.float64, copy=True).reshape(3, -1)
        n = pos.shape[1]
        x, y, z = [np.ascontiguousarray(pos[k]) for k in range(3)]
        fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
        st = np.zeros(6) if stress else None
        t = np.ascontiguousarray(types, dtype=np.int32)
        q = _f64(charges) if charges is not None else None
        e = Energy()
        check(load().akmc_field_forces(self._h, _dp(x), _dp(y), _dp(z), _dp(fx), _dp(fy), _dp(fz), _dp(st),
                                       _dp(_f64(lat)), _dp(q), _ip(t), None, n, C.byref(e)))
        return dict(pos=np.stack([x, y, z]), force=np.stack([fx, fy, fz]), stress=st, energy=e.as_dict())

    def calculate_energy(self, pos, types, lat, charges=None):
        """Field::calculateEnergy (akmc_field_energy): the energy record only."""
        pos = np.array(pos, dtype=np.float64, copy=True).reshape(3, -1)
        n = pos.shape[1]
        x, y, z = [np.ascontiguousarray(pos[k]) for k in range(3)]
        t = np.ascontiguousarray(types, dtype=np.int32)
        q = _f64(charges) if charges is not None else None
        e = Energy()
        check(load().akmc_field_energy(self._h, _dp(x), _dp(y), _dp(z), _dp(_f64(lat)), _dp(q), _ip(t), n,
                                       C.byref(e)))
        return e.as_dict()


class Batch:
    """W walker cells resident in HBM."""

    def __init__(self, field: Field, cells):
        """cells: sequence of objects with .pos (3,N), .types, .lat, .charges."""
        L = load()
        self.field = field
        self.natoms = np.array([c.pos.shape[1] for c in cells], dtype=np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(self.natoms)]).astype(np.int64)
        lat = _f64(np.concatenate([np.asarray(c.lat, dtype=np.float64).reshape(9) for c in cells]))
        h = C.c_void_p()
        check(L.akmc_batch_create(field.handle, len(cells), _ip(self.natoms), _dp(lat), C.byref(h)))
        self._h = h
        self.W = len(cells)
        self.A = int(self.offsets[-1])
        self.upload(cells)

    def upload(self, cells):
        pos = np.concatenate([np.asarray(c.pos, dtype=np.float64) for c in cells], axis=1)
        t = np.ascontiguousarray(np.concatenate([c.types for c in cells]), dtype=np.int32)
        q = None
        if cells[0].charges is not None:
            q = _f64(np.concatenate([c.charges for c in cells]))
        x, y, z = [np.ascontiguousarray(pos[k]) for k in range(3)]
        check(load().akmc_batch_upload(self._h, _dp(x), _dp(y), _dp(z), _ip(t), _dp(q), None))

    def forces(self):
        e = (Energy * self.W)()
        check(load().akmc_batch_forces(self._h, e))
        return [e[w].as_dict() for w in range(self.W)]

    def download(self):
        out = [np.zeros(self.A) for _ in range(6)]
        check(load().akmc_batch_download(self._h, *[_dp(a) for a in out]))
        pos, frc = np.stack(out[:3]), np.stack(out[3:])
        return ([pos[:, self.offsets[w]:self.offsets[w + 1]] for w in range(self.W)],
                [frc[:, self.offsets[w]:self.offsets[w + 1]] for w in range(self.W)])

    def forces_repeat(self, nrep):
        ms = C.c_double()
        check(load().akmc_batch_forces_repeat(self._h, int(nrep), C.byref(ms)))
        return ms.value

    # ---- searches
    def _flat3(self, per_walker):
        """list of (3, N_w) arrays -> [3][A] contiguous FP64"""
        return _f64(np.concatenate([np.asarray(a, dtype=np.float64) for a in per_walker], axis=1).reshape(-1))

    def _mask(self, walkers):
        if walkers is None:
            return None
        m = np.zeros(self.W, dtype=np.int32)
        m[np.asarray(walkers, dtype=int)] = 1
        return m

    def set_frozen(self, frozen_per_walker):
        fr = np.ascontiguousarray(np.concatenate([np.asarray(f) for f in frozen_per_walker]), dtype=np.int32)
        check(load().akmc_batch_set_frozen(self._h, _ip(fr)))

    def start_relax(self, params: RelaxParams | None = None, walkers=None):
        p = params or relax_defaults()
        m = self._mask(walkers)
        check(load().akmc_batch_start_relax(self._h, _ip(m), C.byref(p)))

    def start_dimer(self, mode_per_walker, disp_per_walker=None, params: DimerParams | None = None, walkers=None,
                    x0_per_walker=None, frozen_per_walker=None):
        """Start dimer searches.  Either x0 (absolute start positions) or
        disp (added to the current positions, downloaded first) may be given;
        per-walker lists cover every walker of the batch."""
        p = params or dimer_defaults()
        m = self._mask(walkers)
        mode = self._flat3(mode_per_walker)
        x0 = None
        if x0_per_walker is not None:
            x0 = self._flat3(x0_per_walker)
        elif disp_per_walker is not None:
            cur = self.positions()
            x0 = self._flat3([c + d for c, d in zip(cur, disp_per_walker)])
        fr = None
        if frozen_per_walker is not None:
            fr = np.ascontiguousarray(np.concatenate([np.asarray(f) for f in frozen_per_walker]), dtype=np.int32)
        check(load().akmc_batch_start_dimer(self._h, _ip(m), _dp(x0), _dp(mode), _ip(fr), C.byref(p)))

    def _start_args(self, mode_per_walker, disp_per_walker, walkers, x0_per_walker, frozen_per_walker):
        m = self._mask(walkers)
        mode = self._flat3(mode_per_walker)
        x0 = None
        if x0_per_walker is not None:
            x0 = self._flat3(x0_per_walker)
        elif disp_per_walker is not None:
            cur = self.positions()
            x0 = self._flat3([c + d for c, d in zip(cur, disp_per_walker)])
        fr = None
        if frozen_per_walker is not None:
            fr = np.ascontiguousarray(np.concatenate([np.asarray(f) for f in frozen_per_walker]), dtype=np.int32)
        return m, x0, mode, fr

    def start_art(self, mode_per_walker, disp_per_walker=None, params: ARTParams | None = None, walkers=None,
                  x0_per_walker=None, frozen_per_walker=None):
        """ART searches (optim/ART.cpp:17-756) from x0 (or current + disp)
        along the unit start direction mode (the reference's rVector)."""
        p = params or art_defaults()
        m, x0, mode, fr = self._start_args(mode_per_walker, disp_per_walker, walkers, x0_per_walker,
                                           frozen_per_walker)
        check(load().akmc_batch_start_art(self._h, _ip(m), _dp(x0), _dp(mode), _ip(fr), C.byref(p)))

    def walkers_positions(self, walkers):
        """{w: (3, n_w) search positions} for several walkers, one transfer."""
        return self._walkers_out(load().akmc_batch_get_walkers_positions, walkers)

    def set_harvest(self, on=True):
        """every run chunk copies the positions of the walkers that ended in
        it to host memory (read with harvested_positions while the next
        chunk runs)"""
        check(load().akmc_batch_set_harvest(self._h, int(bool(on))))

    def harvested_positions(self, walkers):
        """{w: (3, n_w)} for walkers harvested by the last waited chunk"""
        return self._walkers_out(load().akmc_batch_get_harvested_positions, walkers)

    def _walkers_out(self, fn, walkers):
        ws = np.ascontiguousarray(walkers, dtype=np.int32)
        n = self.natoms[ws]
        buf = np.zeros(int(3 * n.sum()))
        check(fn(self._h, int(ws.size), _ip(ws), _dp(buf)))
        out, o = {}, 0
        for w, k in zip(ws, n):
            out[int(w)] = buf[o:o + 3 * k].reshape(3, k)
            o += 3 * k
        return out

    def status_array(self):
        """Walker statuses as a numpy structured array (fields of WalkerStatus)."""
        st = (WalkerStatus * self.W)()
        check(load().akmc_batch_status(self._h, st))
        return np.frombuffer(st, dtype=_STATUS_DTYPE).copy()

    def walker_positions(self, w):
        n = int(self.natoms[w])
        out = [np.zeros(n) for _ in range(3)]
        check(load().akmc_batch_get_walker_positions(self._h, int(w), *[_dp(a) for a in out]))
        return np.stack(out)

    def start_flat(self, kind, mask, x0, mode, frozen, params):
        """Low-level start on flat [3][A] / [A] arrays (only the flagged
        walkers' slices are read); kind is "dimer" or "art"."""
        fn = load().akmc_batch_start_dimer if kind == "dimer" else load().akmc_batch_start_art
        check(fn(self._h, _ip(mask), _dp(x0), _dp(mode), _ip(frozen), C.byref(params)))

    def set_verify(self, nl=16, delta=0.01):
        check(load().akmc_batch_set_verify(self._h, int(nl), float(delta)))

    def search_energies(self, bits=64):
        e = np.zeros(self.W)
        check(load().akmc_batch_search_energies(self._h, int(bits), _dp(e)))
        return e

    def run(self, max_steps, sync_every=1):
        n = C.c_int()
        check(load().akmc_batch_run(self._h, int(max_steps), int(sync_every), C.byref(n)))
        return n.value

    def run_async(self, steps):
        """enqueue one chunk of micro-steps and return at once"""
        check(load().akmc_batch_run_async(self._h, int(steps)))

    def run_next(self, steps):
        """run_async without the host round trip before the launch (the
        states read back, the active flags written): for a caller that
        changes walkers only through starts between chunks"""
        check(load().akmc_batch_run_next(self._h, int(steps)))

    def run_wait(self):
        """wait for the chunk of run_async: (steps run, walkers still running)"""
        n, r = C.c_int(), C.c_int()
        check(load().akmc_batch_run_wait(self._h, C.byref(n), C.byref(r)))
        return n.value, r.value

    def status(self):
        st = (WalkerStatus * self.W)()
        check(load().akmc_batch_status(self._h, st))
        return [st[w].as_dict() for w in range(self.W)]

    def positions(self):
        out = [np.zeros(self.A) for _ in range(3)]
        check(load().akmc_batch_get_positions(self._h, *[_dp(a) for a in out]))
        pos = np.stack(out)
        return [pos[:, self.offsets[w]:self.offsets[w + 1]] for w in range(self.W)]

    def mode(self):
        m = np.zeros(3 * self.A)
        check(load().akmc_batch_get_mode(self._h, _dp(m)))
        m = m.reshape(3, self.A)
        return [m[:, self.offsets[w]:self.offsets[w + 1]] for w in range(self.W)]

    def time_phases(self, nrep):
        ms = np.zeros(4)
        check(load().akmc_batch_time_phases(self._h, int(nrep), _dp(ms)))
        return ms / nrep

    def set_graphs(self, enable: bool):
        check(load().akmc_batch_set_graphs(self._h, int(bool(enable))))

    def set_logic_block(self, threads: int):
        """workgroup size of the walker state machines (256 or 1024)"""
        check(load().akmc_batch_set_logic_block(self._h, int(threads)))

    def set_search_zone(self, enable: bool):
        """Active zones of the search micro-steps (on by default)."""
        check(load().akmc_batch_set_search_zone(self._h, int(bool(enable))))

    def time_search_phases(self, nrep):
        """per-phase ms of one evaluation of the search walkers as the micro-steps run it"""
        ms = np.zeros(4)
        check(load().akmc_batch_time_search_phases(self._h, int(nrep), _dp(ms)))
        return ms / nrep

    def set_search_variant(self, v=None):
        """(lanes, unroll, staged) of pass 1, of pass 2, then of the mixed steps'
        FP64 passes for the search micro-steps; None resets"""
        a = None
        if v is not None:
            a = np.zeros(9, np.int32)
            a[:len(v)] = v
        check(load().akmc_batch_set_search_variant(self._h, _ip(a)))

    def autotune_search(self, nrep=5):
        """fastest variant per part on the running searches: (best[8],
        {variant: (ms pass 1, ms pass 2, ms FP64 passes or None)}), ms per nrep evaluations"""
        best = np.zeros(9, np.int32)
        n = 32
        m1, m2, m64 = np.zeros(n), np.zeros(n), np.zeros(n)
        check(load().akmc_batch_autotune_search(self._h, int(nrep), _ip(best), _dp(m1), _dp(m2), _dp(m64)))
        table = {(l, u, st): (m1[16 * st + 4 * li + ui], m2[16 * st + 4 * li + ui], m64[16 * st + 4 * li + ui])
                 for st in (0, 1) for li, l in enumerate((1, 2, 4, 8)) for ui, u in enumerate((1, 2, 4, 8))
                 if m1[16 * st + 4 * li + ui] >= 0}
        return tuple(int(x) for x in best), table

    def zone(self):
        """(member [A], size [W], flags [W], active [W]) of the active zones"""
        member, size = np.zeros(self.A, np.int32), np.zeros(self.W, np.int32)
        flags, active = np.zeros(self.W, np.int32), np.zeros(self.W, np.int32)
        check(load().akmc_batch_get_zone(self._h, _ip(member), _ip(size), _ip(flags), _ip(active)))
        return member, size, flags, active

    def kernel_timer(self, enable: bool):
        check(load().akmc_batch_kernel_timer(self._h, int(bool(enable))))

    def kernel_time(self):
        """(total ms, launches) of the FP32 Sutton-Chen pass-2 kernel timed
        inside run() since the timer was started."""
        ms = C.c_double()
        n = C.c_longlong()
        check(load().akmc_batch_kernel_time(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def eval_counts(self):
        """FP32 force evaluations consumed per walker (int64 [W])."""
        c = np.zeros(self.W, dtype=np.int64)
        check(load().akmc_batch_eval_counts(self._h, c.ctypes.data_as(C.POINTER(C.c_longlong))))
        return c

    def set_variant(self, lanes, unroll=1):
        check(load().akmc_batch_set_variant(self._h, int(lanes), int(unroll)))

    def set_staging(self, on=True):
        """LDS staging of the Sutton-Chen pair kernels (raises when the
        largest walker does not fit one workgroup's LDS)."""
        check(load().akmc_batch_set_staging(self._h, 1 if on else 0))

    def staging(self):
        """(staged now, staging possible)"""
        on, ok = C.c_int(0), C.c_int(0)
        check(load().akmc_batch_staging(self._h, C.byref(on), C.byref(ok)))
        return bool(on.value), bool(ok.value)

    def autotune(self, nrep=10):
        """Time the Sutton-Chen kernel variants (lanes per atom 1/2/4/8 x
        gathers in flight 1/2/4/8 x LDS staging off/on where possible) on
        the resident positions and keep the fastest; returns ((lanes, unroll,
        staged), {(lanes, unroll, staged): ms per force evaluation})."""
        ms = np.zeros(32)
        best = (C.c_int * 3)()
        check(load().akmc_batch_autotune(self._h, int(nrep), best, _dp(ms)))
        table = {(l, u, st): ms[16 * st + 4 * li + ui] / nrep for st in (0, 1) for li, l in enumerate((1, 2, 4, 8))
                 for ui, u in enumerate((1, 2, 4, 8)) if ms[16 * st + 4 * li + ui] >= 0}
        return (best[0], best[1], best[2]), table

    def save_basin(self):
        check(load().akmc_batch_save_basin(self._h))

    def set_basins(self, mask=None, x=None):
        """basins of the flagged walkers (mask: int32 [W], None = all) from the
        flat [3][A] host array x, or from their current positions when x is None"""
        check(load().akmc_batch_set_basins(self._h, _ip(mask), _dp(x)))

    def set_basin2(self, params: Basin2Params | None):
        """basin-2 stage after every converged search (None: off)"""
        check(load().akmc_batch_set_basin2(self._h, C.byref(params) if params is not None else None))

    def walkers_saddles(self, walkers):
        """{w: (3, n_w) saddle positions kept by the basin-2 stage}"""
        return self._walkers_out(load().akmc_batch_get_walkers_saddles, walkers)

    def start_relax_at(self, mask, x0=None, frozen=None, params: RelaxParams | None = None):
        """relax the flagged walkers from flat [3][A] positions x0 with the
        flat [A] frozen mask (None: keep current ones)"""
        p = params or relax_defaults()
        check(load().akmc_batch_start_relax_at(self._h, _ip(mask), _dp(x0), _ip(frozen), C.byref(p)))

    def cancel(self, walkers=None):
        check(load().akmc_batch_cancel(self._h, _ip(self._mask(walkers))))

    def restore_basin(self, walkers=None):
        check(load().akmc_batch_restore_basin(self._h, _ip(self._mask(walkers))))

    @property
    def rebuilds(self):
        return load().akmc_batch_rebuilds(self._h)

    def close(self):
        if getattr(self, "_h", None):
            load().akmc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sc_potentials_text(species_pots, rc, species=None, noimage=False):
    """Reference 'potentials' text for Sutton-Chen (field order
    eps a n m c, potentials/metal/ManyBody.cpp:148-163)."""
    species = species or sorted({p[0] for p in species_pots} | {p[1] for p in species_pots})
    lines = [f"cutoff {rc!r}", f"species {len(species)}"]
    lines += [f"{s} 63.546 0.0 29" for s in species]
    if noimage:
        lines.append("noimage")
    lines.append(f"manybody {len(species_pots)} ev")
    for (A, B, eps, a, n, m, c) in species_pots:
        lines += ["suttonchen", f"{A} {B} {eps!r} {a!r} {n!r} {m!r} {c!r}"]
    lines.append("close")
    return "\n".join(lines) + "\n"


def eam_potentials_text(entries, rc, species=None, noimage=False):
    """Reference 'potentials' text for tabulated EAM: entries are
    (A, B, setfl file path) triples, read as LAMMPS tables
    (potentials/metal/ManyBody.cpp:269-307)."""
    species = species or sorted({e[0] for e in entries} | {e[1] for e in entries})
    lines = [f"cutoff {rc!r}", f"species {len(species)}"]
    lines += [f"{s} 63.546 0.0 29" for s in species]
    if noimage:
        lines.append("noimage")
    lines.append(f"manybody {len(entries)} ev")
    for (A, B, path) in entries:
        lines += ["eam", f"{A} {B} {path} lammps"]
    lines.append("close")
    return "\n".join(lines) + "\n"


def buck_potentials_text(species, charges, pots, rc, noimage=False, ewald_precis=None, ewald_damp=None):
    """Reference 'potentials' text for Buckingham + Ewald
    (potentials/pair/Field.cpp:215-317, TwoBody.cpp:56-112)."""
    lines = [f"cutoff {rc!r}", f"species {len(species)}"]
    lines += [f"{s} 1.0 {q!r}" for s, q in zip(species, charges)]
    if noimage:
        lines.append("noimage")
    if ewald_precis is not None:
        lines.append(f"ewald precis {ewald_precis!r}")
    if ewald_damp is not None:
        lines.append(f"ewald damp {ewald_damp!r}")   # damped-shifted Coulomb (pair Field.cpp:265-270)
    lines.append(f"twobody {len(pots)} ev ev")
    for pot in pots:
        # (A, B, A_buck, rho, C) or (form, A, B, params...) with form one of
        # buck, morse, ljones, bhm, severnsix (TwoBody.cpp:98-245)
        if pot[0] in ("buck", "morse", "ljones", "bhm", "severnsix"):
            form, rest = pot[0], pot[1:]
        else:
            form, rest = "buck", pot
        lines += [form, f"{rest[0]} {rest[1]} " + " ".join(repr(float(v)) for v in rest[2:])]
    lines.append("close")
    return "\n".join(lines) + "\n"


# ---------------------------------------------------------------- text formats
# (host only; see include/akmc.h for the reference formats each follows)

def _b(s):
    return str(s).encode()


def _soa(pos):
    return _f64(np.asarray(pos, dtype=np.float64).reshape(3, -1).reshape(-1))


def dump_basis(path, title, species, labels, pos, lat, time, energy, iteration, header=True, append=False):
    """Basis::dumpBasis: the basis / archive format (species: list of names)."""
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    check(load().akmc_io_dump_basis(_b(path), int(append), int(header), _b(title), _b(" ".join(species)),
                                    int(lab.size), _ip(lab), _dp(_soa(pos)), _dp(_f64(lat)), float(time),
                                    float(energy), int(iteration)))


def write_restart(path, title, species, labels, pos, lat, time, energy, iteration, header=False):
    """runKMC's restart file: 'restart title', ' 0 3 N N', then dumpBasis."""
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    check(load().akmc_io_write_restart(_b(path), int(header), _b(title), _b(" ".join(species)), int(lab.size),
                                       _ip(lab), _dp(_soa(pos)), _dp(_f64(lat)), float(time), float(energy),
                                       int(iteration)))


def read_basis(path, species):
    """Basis::inputBasis: dict(title, labels, pos (3, n), lat (9,), iteration, time, levcfg)."""
    L = load()
    n, lev = C.c_int(), C.c_int()
    check(L.akmc_io_basis_size(_b(path), C.byref(n), C.byref(lev)))
    lab = np.zeros(n.value, dtype=np.int32)
    pos = np.zeros(3 * n.value)
    lat = np.zeros(9)
    it, t = C.c_int(), C.c_double()
    title = C.create_string_buffer(4096)
    check(L.akmc_io_read_basis(_b(path), _b(" ".join(species)), n.value, _ip(lab), _dp(pos), _dp(lat),
                               C.byref(it), C.byref(t), title, 4096))
    return dict(title=title.value.decode(), labels=lab, pos=pos.reshape(3, -1), lat=lat, iteration=it.value,
                time=t.value, levcfg=lev.value)


def write_events(path, species, labels, basin1, lat, basin1_energy, events, chosen, time, fp32=True):
    """writeEventList: events is a list of dicts with act, centre, saddle_energy,
    saddle (3, n), basin2_energy, basin2 (3, n)."""
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    n, m = int(lab.size), len(events)
    arr = lambda k: _f64([float(e[k]) for e in events]) if m else None
    cen = _f64(np.array([e["centre"] for e in events], dtype=np.float64).reshape(-1)) if m else None
    sad = _f64(np.stack([np.asarray(e["saddle"], dtype=np.float64) for e in events]).reshape(-1)) if m else None
    b2 = _f64(np.stack([np.asarray(e["basin2"], dtype=np.float64) for e in events]).reshape(-1)) if m else None
    check(load().akmc_io_write_events(_b(path), _b(" ".join(species)), n, _ip(lab), _dp(_soa(basin1)),
                                      _dp(_f64(lat)), float(basin1_energy), m, _dp(arr("act")), _dp(cen),
                                      _dp(arr("saddle_energy")), _dp(sad), _dp(arr("basin2_energy")), _dp(b2),
                                      int(chosen), float(time), int(bool(fp32))))


def event_pack(global_no, basin2, saddle, e1, e2, es, ea, centre):
    """sendEvent's packet: (int32 [n], float32 [6n + 7])."""
    g = np.ascontiguousarray(global_no, dtype=np.int32)
    n = int(g.size)
    veci = np.zeros(n, dtype=np.int32)
    vecf = np.zeros(6 * n + 7, dtype=np.float32)
    check(load().akmc_kmc_event_pack(n, _ip(g), _dp(_soa(basin2)), _dp(_soa(saddle)), float(e1), float(e2),
                                     float(es), float(ea), _dp(_f64(centre)), _ip(veci),
                                     vecf.ctypes.data_as(C.POINTER(C.c_float))))
    return veci, vecf


def event_unpack(veci, vecf, prefactor, temperature):
    """Event(int*, float*, natoms, prefactor, temperature)."""
    vi = np.ascontiguousarray(veci, dtype=np.int32)
    vf = np.ascontiguousarray(vecf, dtype=np.float32)
    n = int(vi.size)
    g = np.zeros(n, dtype=np.int32)
    b2, sd = np.zeros(3 * n), np.zeros(3 * n)
    en, cen, rate = np.zeros(4), np.zeros(3), C.c_double()
    check(load().akmc_kmc_event_unpack(n, _ip(vi), vf.ctypes.data_as(C.POINTER(C.c_float)), float(prefactor),
                                       float(temperature), _ip(g), _dp(b2), _dp(sd), _dp(en), _dp(cen),
                                       C.byref(rate)))
    return dict(global_no=g, basin2=b2.reshape(3, -1), saddle=sd.reshape(3, -1), basin1_energy=en[0],
                basin2_energy=en[1], saddle_energy=en[2], activation_energy=en[3], centre=cen, rate=rate.value)
