"""ctypes front end of the CPU FP64 oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It wraps oracle/libakmc_oracle.so (strict IEEE build) or, for
the timed CPU baseline, oracle/libakmc_oracle_fast.so (-Ofast -fopenmp).
Arrays follow the reference's Field conventions: SoA positions (x, y, z),
lattice vectors as rows of ``lat``, ``rcp`` its inverse.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAXSPEC = 8
MAXPOT = 36
CTOEV = 14.3997584
EVTODL = 9648.530821

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


class SCField(C.Structure):
    _fields_ = [("nspec", C.c_int), ("npot", C.c_int),
                ("table", C.c_int * (MAXSPEC * MAXSPEC)),
                ("par", C.c_double * (MAXPOT * 9)),
                ("rc", C.c_double), ("min_image", C.c_int)]


class IonField(C.Structure):
    _fields_ = [("nspec", C.c_int), ("npot", C.c_int),
                ("table", C.c_int * (MAXSPEC * MAXSPEC)),
                ("par", C.c_double * (MAXPOT * 6)),
                ("rc", C.c_double), ("min_image", C.c_int), ("do_ewald", C.c_int),
                ("eta", C.c_double), ("cfact", C.c_double), ("cells", C.c_int * 3),
                ("nk", C.c_int), ("gx", _D), ("gy", _D), ("gz", _D), ("g0", _D), ("g1", _D),
                ("form", C.c_int * MAXPOT), ("energy_path", C.c_int),
                ("dsf_e", _D), ("dsf_f", _D), ("dsf_del", C.c_double), ("dsf_rfld0", C.c_double),
                ("dsf_rfld1", C.c_double)]

# points of the damped-shifted Coulomb mesh: the reference's MAXMESH, a
# constant of its missing Constants.h; this value is the build's choice
DSF_MESH = 5000


def set_damped(f, alpha, rc, nmesh=DSF_MESH):
    """switch an ion field to damped-shifted Coulomb (ewald damp alpha,
    pair Field.cpp:119-126): builds the erfc mesh of fieldErfcGen"""
    e, fm = np.zeros(nmesh), np.zeros(nmesh)
    r0, r1 = C.c_double(), C.c_double()
    lib().ako_damped_mesh.restype = C.c_double
    dl = lib().ako_damped_mesh(C.c_double(alpha), C.c_double(rc), nmesh, _p(e), _p(fm), C.byref(r0), C.byref(r1))
    f.do_ewald = 2
    f.eta = alpha
    f.dsf_e, f.dsf_f = _p(e), _p(fm)
    f.dsf_del, f.dsf_rfld0, f.dsf_rfld1 = dl, r0.value, r1.value
    f._dsf = (e, fm)  # keep alive
    return e, fm

# short-range pair forms of the potentials text (TwoBody.cpp:440-470) and
# their parameter counts
PAIR_FORMS = {"buck": (0, 3), "morse": (1, 4), "ljones": (2, 2), "bhm": (3, 5), "severnsix": (4, 2)}


class FieldCtx(C.Structure):
    _fields_ = [("field", C.c_void_p), ("kind", C.c_int), ("type", _I), ("q", _D),
                ("lat", _D), ("rcp", _D), ("ncalls", C.c_longlong)]


class RelaxCntrl(C.Structure):
    _fields_ = [("maxIter", C.c_int), ("normTol", C.c_double), ("forceTol", C.c_double),
                ("N_min", C.c_int), ("initTimeStep", C.c_double), ("alphaDecrease", C.c_double),
                ("stepIncrease", C.c_double), ("stepDecrease", C.c_double),
                ("alphaStart", C.c_double), ("maxTimeStep", C.c_double), ("euler", C.c_int),
                ("timeStepUnit", C.c_double), ("N_delaySteps", C.c_int), ("N_maxUphill", C.c_int),
                ("initialDelay", C.c_int), ("minTimeStep", C.c_double)]


class ARTCntrl(C.Structure):
    _fields_ = [("numLanczosVectors", C.c_int), ("deltaX", C.c_double),
                ("initialDisplacement", C.c_double), ("maxEigenvalue", C.c_double),
                ("eigValTol", C.c_double), ("parDamp", C.c_double), ("goPerpendicular", C.c_int),
                ("maxParIter", C.c_int), ("maxPerpIter", C.c_int), ("normTol", C.c_double),
                ("forceTol", C.c_double), ("N_min", C.c_int), ("initTimeStep", C.c_double),
                ("alphaDecrease", C.c_double), ("stepIncrease", C.c_double),
                ("stepDecrease", C.c_double), ("alphaStart", C.c_double),
                ("maxTimeStep", C.c_double), ("initTimeStepPerp", C.c_double),
                ("maxTimeStepPerp", C.c_double), ("timeStepUnit", C.c_double)]


class DimerCntrl(C.Structure):
    _fields_ = [("deltaTheta", C.c_double), ("dimerDisplacement", C.c_double),
                ("maxDeltaX", C.c_double), ("maxRotations", C.c_int), ("maxParIter", C.c_int),
                ("maxEigenvalue", C.c_double), ("normTol", C.c_double), ("forceTol", C.c_double),
                ("N_min", C.c_int), ("alphaDecrease", C.c_double), ("stepIncrease", C.c_double),
                ("stepDecrease", C.c_double), ("alphaStart", C.c_double),
                ("initTimeStep", C.c_double), ("maxTimeStep", C.c_double),
                ("maxStep", C.c_double), ("timeStepUnit", C.c_double)]


FORCE_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_int, _D, _D, _D, _D, _D, _D)

_libs: dict[str, C.CDLL] = {}


def build() -> None:
    """Compile both oracle builds with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(fast: bool = False) -> C.CDLL:
    name = "libakmc_oracle_fast.so" if fast else "libakmc_oracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    L.ako_sc_forces.restype = C.c_double
    L.ako_sc_forces_omp.restype = C.c_double
    L.ako_sc_energy.restype = C.c_double
    L.ako_sc_count_pairs.restype = C.c_longlong
    L.ako_ion_forces.restype = C.c_double
    L.ako_ion_count_pairs.restype = C.c_longlong
    L.ako_ewald_recip_force.restype = C.c_double
    L.ako_ewald_recip_energy.restype = C.c_double
    L.ako_field_cb.restype = C.c_double
    _libs[name] = L
    return L


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def cell(L):
    """Row-vector lattice for an orthorhombic box of edges L (scalar or 3)."""
    L = np.broadcast_to(np.asarray(L, dtype=np.float64), (3,))
    return np.diag(L).reshape(9).copy()


def inverse_cell(lat):
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    rcp = np.zeros(9)
    lib().ako_inverse_cell(_p(lat), _p(rcp))
    return rcp


def dcell(lat):
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    out = np.zeros(10)
    lib().ako_dcell(_p(lat), _p(out))
    return out


def sc_field(species, pots, rc, min_image=True):
    """species: list of names; pots: list of (A, B, eps, a, n, m, c)."""
    f = SCField()
    f.nspec = len(species)
    f.npot = len(pots)
    for i in range(MAXSPEC * MAXSPEC):
        f.table[i] = -1
    for k, (A, B, eps, a, n, m, c) in enumerate(pots):
        for i, si in enumerate(species):
            for j, sj in enumerate(species):
                if si == A and sj == B:
                    f.table[j * f.nspec + i] = k
                    f.table[i * f.nspec + j] = k
        for t, v in enumerate((eps, a, n, m, c)):
            f.par[9 * k + t] = v
    f.rc = rc
    f.min_image = 1 if min_image else 0
    return f


def sc_forces(f, pos, types, lat, omp=False, fast=False):
    """Returns dict(pos=wrapped, force (3,N), stress, e_many, e_pair, energy, rho)."""
    pos = np.array(pos, dtype=np.float64, copy=True).reshape(3, -1)
    n = pos.shape[1]
    x, y, z = [np.ascontiguousarray(pos[k]) for k in range(3)]
    types = np.ascontiguousarray(types, dtype=np.int32)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    rcp = inverse_cell(lat)
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    st = np.zeros(6)
    em, ep = C.c_double(), C.c_double()
    rho = np.zeros(n)
    L = lib(fast)
    if omp:
        e = L.ako_sc_forces_omp(C.byref(f), n, _p(x), _p(y), _p(z), _p(types, C.c_int), _p(lat), _p(rcp),
                                _p(fx), _p(fy), _p(fz), _p(st), C.byref(em), C.byref(ep))
    else:
        e = L.ako_sc_forces(C.byref(f), n, _p(x), _p(y), _p(z), _p(types, C.c_int), _p(lat), _p(rcp),
                            _p(fx), _p(fy), _p(fz), _p(st), C.byref(em), C.byref(ep), _p(rho))
    return dict(pos=np.stack([x, y, z]), force=np.stack([fx, fy, fz]), stress=st,
                e_many=em.value, e_pair=ep.value, energy=e, rho=rho)


def sc_energy(f, pos, types, lat):
    pos = np.array(pos, dtype=np.float64, copy=True).reshape(3, -1)
    n = pos.shape[1]
    x, y, z = [np.ascontiguousarray(pos[k]) for k in range(3)]
    types = np.ascontiguousarray(types, dtype=np.int32)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    rcp = inverse_cell(lat)
    em, ep = C.c_double(), C.c_double()
    e = lib().ako_sc_energy(C.byref(f), n, _p(x), _p(y), _p(z), _p(types, C.c_int), _p(lat), _p(rcp),
                            C.byref(em), C.byref(ep))
    return e, em.value, ep.value


def sc_count_pairs(f, pos, lat):
    pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(3, -1))
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    rcp = inverse_cell(lat)
    return lib().ako_sc_count_pairs(C.byref(f), pos.shape[1], _p(pos[0].copy()), _p(pos[1].copy()),
                                    _p(pos[2].copy()), _p(lat), _p(rcp))


class Ewald:
    """Holds the k-vector arrays so the IonField pointers stay valid."""

    def __init__(self, lat, eta, gcell=None, acc=None, rc=None, cfact=CTOEV, maxk=200000):
        lat = np.ascontiguousarray(lat, dtype=np.float64)
        rcp = inverse_cell(lat)
        cp = dcell(lat)
        L = lib()
        if gcell is None:
            e_eta, eps, tol = C.c_double(), C.c_double(), C.c_double()
            L.ako_ewald_eta_from_acc(C.c_double(acc), C.c_double(rc), C.byref(e_eta), C.byref(eps), C.byref(tol))
            if eta is None:
                eta = e_eta.value
            gc = (C.c_int * 3)()
            cut = C.c_double()
            L.ako_ewald_gcell(_p(lat), _p(rcp), C.c_double(eta), C.c_double(rc), eps, tol, gc, C.byref(cut))
            gcell = [gc[0], gc[1], gc[2]]
            rcpcut = cut.value
        else:
            gcell = list(np.broadcast_to(np.asarray(gcell, dtype=int), (3,)))
            rcp_cp = dcell(rcp)
            rcpcut = min(gcell[k] * rcp_cp[6 + k] for k in range(3))
        self.eta, self.gcell, self.rcpcut = eta, gcell, rcpcut
        bufs = [np.zeros(maxk) for _ in range(5)]
        gci = (C.c_int * 3)(*[int(g) for g in gcell])
        nk = L.ako_ewald_kvectors(_p(rcp), C.c_double(cp[9]), C.c_double(eta), C.c_double(cfact), gci,
                                  C.c_double(rcpcut), maxk, *[_p(b) for b in bufs])
        if nk < 0:
            raise ValueError("too many k vectors")
        self.nk = nk
        self.gx, self.gy, self.gz, self.g0, self.g1 = [np.ascontiguousarray(b[:nk]) for b in bufs]


def ion_field(species, pots, rc, ewald: Ewald | None, min_image=True, lat=None, cfact=CTOEV,
              table_default=0):
    """pots: list of (A, B, Abuck, rho, C) Buckingham entries, or of
    (form, A, B, params...) with form a key of PAIR_FORMS.  table_default=0
    reproduces the reference's pair-table initialisation (TwoBody.cpp:509)."""
    f = IonField()
    f.nspec = len(species)
    f.npot = len(pots)
    for i in range(MAXSPEC * MAXSPEC):
        f.table[i] = table_default
    for k, pot in enumerate(pots):
        if isinstance(pot[0], str) and pot[0] in PAIR_FORMS:
            form, npar = PAIR_FORMS[pot[0]]
            A, B, par = pot[1], pot[2], pot[3:3 + npar]
        else:
            form, (A, B, *par) = 0, pot
        for i, si in enumerate(species):
            for j, sj in enumerate(species):
                if si == A and sj == B:
                    f.table[j * f.nspec + i] = k
                    f.table[i * f.nspec + j] = k
        for t, v in enumerate(par):
            f.par[6 * k + t] = v
        f.form[k] = form
    f.rc = rc
    f.min_image = 1 if min_image else 0
    f.cfact = cfact
    if lat is not None:
        cp = dcell(np.asarray(lat, dtype=np.float64))
        for k in range(3):
            f.cells[k] = int(rc / cp[6 + k]) + 1
    if ewald is not None:
        f.do_ewald = 1
        f.eta = ewald.eta
        f.nk = ewald.nk
        f.gx, f.gy, f.gz, f.g0, f.g1 = [_p(a) for a in (ewald.gx, ewald.gy, ewald.gz, ewald.g0, ewald.g1)]
        f._ewald = ewald  # keep alive
    else:
        f.do_ewald = 0
    return f


def ion_forces(f, pos, types, charges, lat):
    pos = np.array(pos, dtype=np.float64, copy=True).reshape(3, -1)
    n = pos.shape[1]
    x, y, z = [np.ascontiguousarray(pos[k]) for k in range(3)]
    types = np.ascontiguousarray(types, dtype=np.int32)
    q = np.ascontiguousarray(charges, dtype=np.float64)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    rcp = inverse_cell(lat)
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    st = np.zeros(6)
    er, ee, ev = C.c_double(), C.c_double(), C.c_double()
    e = lib().ako_ion_forces(C.byref(f), n, _p(x), _p(y), _p(z), _p(types, C.c_int), _p(q), _p(lat),
                             _p(rcp), _p(fx), _p(fy), _p(fz), _p(st), C.byref(er), C.byref(ee), C.byref(ev))
    return dict(pos=np.stack([x, y, z]), force=np.stack([fx, fy, fz]), stress=st,
                e_rcp=er.value, e_real=ee.value, e_vdw=ev.value, energy=e)


def ion_count_pairs(f, pos, lat):
    pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(3, -1))
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    rcp = inverse_cell(lat)
    return lib().ako_ion_count_pairs(C.byref(f), pos.shape[1], _p(pos[0].copy()), _p(pos[1].copy()),
                                     _p(pos[2].copy()), _p(lat), _p(rcp))


def tridiag_eig(d, e):
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    n = d.size
    w = np.zeros(n)
    z = np.zeros(n * n)
    rc = lib().ako_tridiag_eig(n, _p(d), _p(e if e.size else np.zeros(1)), _p(w), _p(z))
    if rc != 0:
        raise RuntimeError("tridiagonal eigensolver did not converge")
    return w, z.reshape(n, n).T  # columns are eigenvectors


class FieldHandle:
    """Binds a field + configuration into the C force callback."""

    def __init__(self, field, kind, types, lat, charges=None, fast=False):
        self.field = field
        self._fast = fast
        self.types = np.ascontiguousarray(types, dtype=np.int32)
        self.lat = np.ascontiguousarray(lat, dtype=np.float64)
        self.rcp = inverse_cell(self.lat)
        self.q = np.ascontiguousarray(charges if charges is not None else np.zeros(len(types)), dtype=np.float64)
        self.ctx = FieldCtx(C.cast(C.pointer(field), C.c_void_p), kind, _p(self.types, C.c_int), _p(self.q),
                            _p(self.lat), _p(self.rcp), 0)
        self.fn = C.cast(lib(fast).ako_field_cb, FORCE_FN)

    @property
    def ncalls(self):
        return self.ctx.ncalls


def _xyz(pos):
    pos = np.array(pos, dtype=np.float64, copy=True).reshape(3, -1)
    return [np.ascontiguousarray(pos[k]) for k in range(3)]


def relax_defaults():
    c = RelaxCntrl()
    lib().ako_relax_defaults(C.byref(c))
    return c


def art_defaults():
    c = ARTCntrl()
    lib().ako_art_defaults(C.byref(c))
    return c


def dimer_defaults(fast=False):
    c = DimerCntrl()
    lib(fast).ako_dimer_defaults(C.byref(c))
    return c


def fire_relax(h: FieldHandle, pos, frozen=None, cntrl=None, method="fire"):
    """Relax::FIRERelax (method "fire") or Relax::FIRE2Relax ("fire2");
    cntrl.euler = 0 is the reference's verletIntegration."""
    x, y, z = _xyz(pos)
    n = x.size
    fr = np.ascontiguousarray(frozen if frozen is not None else np.zeros(n), dtype=np.int32)
    c = cntrl or relax_defaults()
    e = C.c_double()
    it = C.c_int()
    fn = {"fire": lib().ako_fire_relax, "fire2": lib().ako_fire2_relax}[method]
    st = fn(h.fn, C.byref(h.ctx), n, _p(x), _p(y), _p(z), _p(fr, C.c_int), C.byref(c), C.byref(e), C.byref(it))
    return dict(status=st, pos=np.stack([x, y, z]), energy=e.value, iters=it.value)


def active_sphere(pos, lat, centre, radius, frozen=None):
    x, y, z = _xyz(pos)
    n = x.size
    fr = np.ascontiguousarray(frozen if frozen is not None else np.zeros(n), dtype=np.int32).copy()
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    rcp = inverse_cell(lat)
    act = lib().ako_active_sphere(n, _p(x), _p(y), _p(z), _p(lat), _p(rcp), int(centre), C.c_double(radius),
                                  _p(fr, C.c_int))
    return fr, act


def random_displacement(pos, frozen, u, d0):
    x, y, z = _xyz(pos)
    fr = np.ascontiguousarray(frozen, dtype=np.int32)
    u = np.ascontiguousarray(u, dtype=np.float64)
    rv = np.zeros(u.size)
    lib().ako_random_displacement(x.size, _p(x), _p(y), _p(z), _p(fr, C.c_int), _p(u), C.c_double(d0), _p(rv))
    return np.stack([x, y, z]), rv


def lanczos(h: FieldHandle, pos, frozen, rvec, nl=20, delta=0.01, tol=1e-3):
    x, y, z = _xyz(pos)
    fr = np.ascontiguousarray(frozen, dtype=np.int32)
    rv = np.ascontiguousarray(rvec, dtype=np.float64).copy()
    m = rv.size
    fpar, fperp = np.zeros(m), np.zeros(m)
    ev, e = C.c_double(), C.c_double()
    conv = lib().ako_lanczos(h.fn, C.byref(h.ctx), x.size, _p(x), _p(y), _p(z), _p(fr, C.c_int), _p(rv), nl,
                             C.c_double(delta), C.c_double(tol), C.byref(ev), _p(fpar), _p(fperp), C.byref(e))
    return dict(converged=conv, eigval=ev.value, mode=rv, fpar=fpar, fperp=fperp, energy=e.value)


def art_search(h: FieldHandle, pos, frozen, rvec, cntrl=None):
    x, y, z = _xyz(pos)
    fr = np.ascontiguousarray(frozen, dtype=np.int32)
    rv = np.ascontiguousarray(rvec, dtype=np.float64).copy()
    c = cntrl or art_defaults()
    e, ev, it = C.c_double(), C.c_double(), C.c_int()
    st = lib().ako_art_search(h.fn, C.byref(h.ctx), x.size, _p(x), _p(y), _p(z), _p(fr, C.c_int), _p(rv),
                              C.byref(c), C.byref(e), C.byref(ev), C.byref(it))
    return dict(status=st, pos=np.stack([x, y, z]), energy=e.value, eigval=ev.value, iters=it.value, mode=rv)


def dimer_search(h: FieldHandle, pos, frozen, mode, cntrl=None, fast=False):
    x, y, z = _xyz(pos)
    fr = np.ascontiguousarray(frozen, dtype=np.int32)
    md = np.ascontiguousarray(mode, dtype=np.float64).copy()
    c = cntrl or dimer_defaults()
    e, cv, it, nf = C.c_double(), C.c_double(), C.c_int(), C.c_int()
    st = lib(fast).ako_dimer_search(h.fn, C.byref(h.ctx), x.size, _p(x), _p(y), _p(z), _p(fr, C.c_int), _p(md),
                                C.byref(c), C.byref(e), C.byref(cv), C.byref(it), C.byref(nf))
    return dict(status=st, pos=np.stack([x, y, z]), energy=e.value, curvature=cv.value, iters=it.value,
                fcalls=nf.value, mode=md)


class PyFieldHandle:
    """A force field written in Python (e.g. an analytic test surface),
    exposed through the same C callback type as FieldHandle."""

    def __init__(self, energy_forces):
        self._ef = energy_forces
        self.ncalls = 0

        def cb(ctx, n, x, y, z, fx, fy, fz):
            self.ncalls += 1
            X = np.ctypeslib.as_array(x, (n,))
            Y = np.ctypeslib.as_array(y, (n,))
            Z = np.ctypeslib.as_array(z, (n,))
            e, F = self._ef(np.stack([X, Y, Z]))
            np.ctypeslib.as_array(fx, (n,))[:] = F[0]
            np.ctypeslib.as_array(fy, (n,))[:] = F[1]
            np.ctypeslib.as_array(fz, (n,))[:] = F[2]
            return float(e)

        self.fn = FORCE_FN(cb)
        self.ctx = C.c_int(0)
