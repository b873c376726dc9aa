"""Adaptive kinetic Monte Carlo on the batched device engine.

``KineticMonteCarlo`` advances S independent KMC trajectories ("systems")
at once.  Each system owns G walkers of one batch resident in HBM: one
walker relaxes the system's basin 1 (kmc/KineticMonteCarlo.cpp:110-115),
then all G run saddle searches from it, each followed on the device by its
FP64 verification and its basin-2 stage (kick, FIRE relax, checkBasin;
checkForNewBasin, :1626-1713).  Walkers whose search fails or lands
outside the activation window are refilled at once, as the reference's
parent refills idle search groups (calculateRateConstant, :352-455); when a
system holds ``events`` valid events its in-flight searches are stopped
(TERMINATED, :446-455) and the KMC step runs on the host with the library's
bookkeeping functions: rates (Event.cpp:71,103), the event choice
(:144-169), the clock (:173, calculateTime :1151-1161), the excluded list
(updateExcludedList :1291-1313), the move (Event::setPositions) and event
recycling (clearEventList :1163-1265).  Systems step independently: one
system's KMC step never waits for another's searches.

A valid event is a converged search whose saddle passes the FP64 tests of
``search.SaddleCampaign`` (largest FP64 force, FP32/FP64 barrier
agreement, one negative curvature, positive barrier), whose basin 2 is new
(checkBasin) and whose FP64 activation energy lies in
[minCap, minCap + window kB T] (kmc/kmc.cpp:160, :1081-1083).  As in the
reference, events are not de-duplicated: an event found twice enters the
rate sum twice.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import lib as L
from . import workloads as W
from .search import active_region


@dataclass
class KMCParams:
    """The reference's kmc{} keys (docs/docs/source/ACDCUserGuide.rst:150-180
    and the KMCCntrl members read in kmc/KineticMonteCarlo.cpp).  Defaults
    follow docs/examples/gdc/control where it sets the key; basin2Delta,
    kmcMaxRadius and the recycling keys live in the reference's missing
    KMCCntrl.h, so theirs are this engine's choice."""
    kmcTemperature: float = 1000.0
    kmcPreFactor: float = 1.0e13
    kmcSteps: int = 2
    kmcImages: int = 5               # kmcevents: valid events per KMC step
    kmcMinCap: float = 0.1
    kmcWindow: float = 50.0
    basin2Delta: float = 0.2
    kmcBasinRadius: float = 0.5
    kmcMaxRadius: float = 5.0
    recycle: bool = False
    recycleRadius: float = 10.0
    recycleType: int = 1             # 2: keep each far event with probability recyclePercentage / 100
    recyclePercentage: float = 100.0
    superBasin: bool = False
    historySize: int = 10

    @property
    def kmcMaxCap(self) -> float:
        return L.load().akmc_kmc_max_cap(self.kmcMinCap, self.kmcWindow, self.kmcTemperature)


@dataclass
class Event:
    """kmc/Event.h: energies, centre, rate, and the basin-2 and saddle
    positions of every atom (kept in FP64)."""
    basin1_energy: float
    basin2_energy: float
    saddle_energy: float
    activation_energy: float
    rate: float
    centre: np.ndarray
    basin2: np.ndarray
    saddle: np.ndarray
    walker: int
    force_calls: int = 0
    basin2_force_calls: int = 0


@dataclass
class KMCStepRecord:
    system: int
    iteration: int
    events: int
    chosen: int
    total_rate: float
    dt: float
    time: float
    activation_energy: float
    basin1_energy: float
    basin2_energy: float
    recycled: int
    barriers: list = field(default_factory=list)   # activation energies of the events in the rate sum


@dataclass
class KMCStats:
    kmc_steps: int = 0
    searches: int = 0
    converged: int = 0
    verified: int = 0
    new_basin: int = 0
    valid_events: int = 0
    out_of_window: int = 0
    force_calls: int = 0          # searches and their verification
    basin2_force_calls: int = 0
    relax_force_calls: int = 0    # basin-1 relaxes
    micro_steps: int = 0
    seconds: float = 0.0
    steps: list = field(default_factory=list)


class _System:
    def __init__(self, cell: W.Cell, walkers, seed):
        self.cell = cell
        self.walkers = list(walkers)
        self.basin = np.array(cell.pos, dtype=np.float64, copy=True)
        self.vacancies = [np.array(v, dtype=np.float64) for v in cell.vacancies]
        self.e_basin32 = 0.0
        self.e_basin64 = 0.0
        self.events: list[Event] = []
        self.excluded = np.zeros(cell.n, dtype=np.int32)
        self.time = 0.0
        self.iteration = 0
        self.phase = "relax"
        self.rng = np.random.default_rng(seed)
        self.site = {}


class KineticMonteCarlo:
    """S systems x G walkers of one device batch; see the module docstring."""

    def __init__(self, potentials_text, systems, walkers_per_system=8, params: KMCParams | None = None,
                 method="dimer", search_params=None, relax=None, region_radius=6.0, d0=0.1, seed=0,
                 fmax_verify=0.05, barrier_tol=0.01, verify_nl=16, verify_delta=0.01, skin=None, sync_every=25,
                 output_dir=None, write_events=False, autotune=True):
        if method not in ("dimer", "art"):
            raise ValueError("the kmc method is not recognised")  # KineticMonteCarlo.cpp:1034-1046
        self.p = params or KMCParams()
        self.G = int(walkers_per_system)
        self.method = method
        self.field = L.Field(potentials_text)
        if skin is not None:
            self.field.set_skin(skin)
        cells = [c for c in systems for _ in range(self.G)]
        self.batch = L.Batch(self.field, cells)
        if autotune:
            # kernel variant of the Sutton-Chen pair kernels timed on the seeded cells
            if self.field.kind == 0:
                self.batch.forces()
                self.batch.autotune(5)
        self.W = len(cells)
        self.A = self.batch.A
        self.off = self.batch.offsets
        self.search_params = search_params or (L.dimer_defaults() if method == "dimer" else L.art_defaults())
        self.relax = relax or L.relax_defaults()
        self.region_radius = region_radius
        self.d0 = d0
        self.fmax_verify = fmax_verify
        self.barrier_tol = barrier_tol
        self.sync_every = sync_every
        self.systems = [_System(c, range(s * self.G, (s + 1) * self.G), seed * 1000003 + s)
                        for s, c in enumerate(systems)]
        self.batch.set_verify(verify_nl, verify_delta)
        self.batch.set_basin2(L.basin2_defaults(basin2Delta=self.p.basin2Delta, kmcBasinRadius=self.p.kmcBasinRadius,
                                                kmcMaxRadius=self.p.kmcMaxRadius))
        self._x0 = np.zeros(3 * self.A)
        self._mode = np.zeros(3 * self.A)
        self._frozen = np.zeros(self.A, dtype=np.int32)
        self.pending = set()
        self._L = L.load()
        self._lat = [np.ascontiguousarray(s.cell.lat, dtype=np.float64) for s in self.systems]
        # files of runKMC (kmc/KineticMonteCarlo.cpp:181-241), one folder per
        # system: events.<iteration> (writeEvents), restart and archive
        self.output_dir = output_dir
        self.write_events = write_events
        if output_dir:
            import os
            for si in range(len(self.systems)):
                os.makedirs(os.path.join(output_dir, f"system{si}"), exist_ok=True)

    # ------------------------------------------------------------ helpers
    def _sys(self, w):
        return self.systems[w // self.G]

    def _put(self, arr3, w, x):
        a, b = self.off[w], self.off[w + 1]
        for c in range(3):
            arr3[c * self.A + a:c * self.A + b] = x[c]

    def _start_relaxes(self, systems):
        mask = np.zeros(self.W, dtype=np.int32)
        for s in systems:
            lead = s.walkers[0]
            self._put(self._x0, lead, s.basin)
            self._frozen[self.off[lead]:self.off[lead + 1]] = 0   # basin1.unFreezeAtoms()
            mask[lead] = 1
            s.phase = "relax"
            self.pending.add(lead)
        self.batch.start_relax_at(mask, self._x0, self._frozen, self.relax)

    def _proposal(self, s: _System):
        v = int(s.rng.integers(0, len(s.vacancies)))
        site = s.vacancies[v]
        frozen = active_region(s.cell, s.basin, site, self.region_radius)
        nb = W.neighbours_of_site(W.Cell(s.basin, s.cell.types, s.cell.lat), site, 12)
        disp, mode = W.local_displacement(W.Cell(s.basin, s.cell.types, s.cell.lat), nb, self.d0,
                                          int(s.rng.integers(0, 2**31 - 1)))
        return s.basin + disp, mode, frozen, site

    def _start_searches(self, walkers):
        if not walkers:
            return
        mask = np.zeros(self.W, dtype=np.int32)
        for w in walkers:
            s = self._sys(w)
            x0, mode, frozen, site = self._proposal(s)
            s.site[w] = site
            self._put(self._x0, w, x0)
            self._put(self._mode, w, mode)
            self._frozen[self.off[w]:self.off[w + 1]] = frozen
            mask[w] = 1
            self.pending.add(w)
        self.batch.start_flat(self.method, mask, self._x0, self._mode, self._frozen, self.search_params)

    # ------------------------------------------------------------ the KMC step
    def _kmc_step(self, s: _System, stats: KMCStats):
        """runKMC's step after the rate constants (kmc/KineticMonteCarlo.cpp:144-205)."""
        Lb = self._L
        lat = self._lat[s.walkers[0] // self.G]
        n = len(s.events)
        rates = np.ascontiguousarray([e.rate for e in s.events], dtype=np.float64)
        total = C.c_double()
        chosen = Lb.akmc_kmc_select(n, rates.ctypes.data_as(C.POINTER(C.c_double)), float(s.rng.random()),
                                    C.byref(total))
        if chosen < 0:
            raise RuntimeError("no event could be chosen (zero total rate)")
        dt = Lb.akmc_kmc_time_increase(total.value, float(s.rng.random()))
        s.time += dt
        ev = s.events[chosen]
        N = s.cell.n
        if self.p.superBasin:
            d = self._displaced(ev.basin2, s.basin, lat)
            ex = s.excluded
            L.check(Lb.akmc_kmc_update_excluded(N, ex.ctypes.data_as(C.POINTER(C.c_int)), len(d),
                                                d.ctypes.data_as(C.POINTER(C.c_int)), int(self.p.historySize)))
        old = s.basin
        s.basin = np.array(ev.basin2, copy=True)            # Event::setPositions
        self._move_vacancy(s, old, s.basin)
        e1 = s.e_basin64
        if self.output_dir and self.write_events:
            self._write_events(s, old, e1, chosen, s.iteration + 1)
        s.e_basin64 = ev.basin2_energy
        kept = self._recycle(s, chosen, lat)
        if self.output_dir:
            self._write_basis(s, s.iteration + 1)
        s.iteration += 1
        stats.kmc_steps += 1
        stats.steps.append(KMCStepRecord(s.walkers[0] // self.G, s.iteration, n, chosen, total.value, dt, s.time,
                                         ev.activation_energy, e1, ev.basin2_energy, len(kept),
                                         [e.activation_energy for e in s.events]))
        s.events = kept

    def _folder(self, s):
        import os
        return os.path.join(self.output_dir, f"system{self.systems.index(s)}")

    def _names(self, s):
        return list(s.cell.species) or ["X"]

    def _write_events(self, s, basin1, e1, chosen, iteration):
        """writeEventList (:1561-1616) before the move; the time written is
        kmcTime x prefactor as at :181"""
        import os
        evs = [dict(act=e.activation_energy, centre=e.centre, saddle_energy=e.saddle_energy, saddle=e.saddle,
                    basin2_energy=e.basin2_energy, basin2=e.basin2) for e in s.events]
        L.write_events(os.path.join(self._folder(s), f"events.{iteration}"), self._names(s), s.cell.types, basin1,
                       s.cell.lat, e1, evs, chosen, s.time * self.p.kmcPreFactor)

    def _write_basis(self, s, iteration):
        """the restart file and the archive after a step (:210-236): basis
        dumps of the new basin 1 with time kmcTime x prefactor; the dump
        header goes to the first file written, as the reference's Basis
        writes it once"""
        import os
        d = self._folder(s)
        first = not getattr(s, "_dumped", False)
        args = ("kmc", self._names(s), s.cell.types, s.basin, s.cell.lat, s.time * self.p.kmcPreFactor,
                s.e_basin64, iteration)
        L.write_restart(os.path.join(d, "restart"), *args, header=first)
        L.dump_basis(os.path.join(d, "archive"), *args, header=False, append=not first)
        s._dumped = True

    def _displaced(self, pos, basin, lat):
        n = pos.shape[1]
        out = np.zeros(n, dtype=np.int32)
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(basin, dtype=np.float64).reshape(-1)
        m = self._L.akmc_kmc_displaced_atoms(n, L._dp(p), L._dp(b), L._dp(lat), float(self.p.kmcBasinRadius),
                                             out.ctypes.data_as(C.POINTER(C.c_int)))
        L.check(m)
        return out[:m]

    def _recycle(self, s: _System, chosen, lat):
        n = len(s.events)
        cen = np.ascontiguousarray([e.centre for e in s.events], dtype=np.float64).reshape(-1)
        u = np.ascontiguousarray(s.rng.random(n) if self.p.recycleType == 2 else np.zeros(n))
        blocked = None
        if self.p.superBasin:
            # the reference compares the excluded list's entries with the
            # event's displaced atom numbers (KineticMonteCarlo.cpp:1228-1234)
            vals = set(int(v) for v in s.excluded)
            blocked = np.ascontiguousarray(
                [int(any(int(a) in vals for a in self._displaced(e.basin2, s.basin, lat))) for e in s.events],
                dtype=np.int32)
        keep = np.zeros(n, dtype=np.int32)
        L.check(self._L.akmc_kmc_recycle(n, L._dp(cen), int(chosen), L._dp(lat), float(self.p.recycleRadius),
                                         int(bool(self.p.recycle)), int(self.p.recycleType),
                                         float(self.p.recyclePercentage), L._dp(u), L._ip(blocked),
                                         keep.ctypes.data_as(C.POINTER(C.c_int))))
        return [e for e, k in zip(s.events, keep) if k]

    @staticmethod
    def _move_vacancy(s: _System, old, new):
        """the proposal sites follow the executed event: the atom that moved
        most leaves its old site empty and fills the vacancy nearest to its
        new position"""
        if not s.vacancies:
            return
        d = W.min_image(new - old, s.cell.lat)
        i = int(np.argmax((d * d).sum(0)))
        dv = [W.min_image(new[:, i] - v, s.cell.lat) for v in s.vacancies]
        k = int(np.argmin([float((x * x).sum()) for x in dv]))
        s.vacancies[k] = np.array(old[:, i], copy=True)

    # ------------------------------------------------------------ harvest
    def _harvest(self, stats: KMCStats):
        st = self.batch.status_array()
        done = [w for w in sorted(self.pending) if st["done"][w]]
        relaxed, searched = [], []
        for w in done:
            self.pending.discard(w)
            (relaxed if st["kind"][w] == 1 else searched).append(w)
        restart_relax, start_search = [], []
        # basin-1 relaxes that ended: every walker of the system gets the basin
        if relaxed:
            pos = self.batch.walkers_positions(relaxed)
            mask = np.zeros(self.W, dtype=np.int32)
            for w in relaxed:
                s = self._sys(w)
                if st["status"][w] != 0:
                    raise RuntimeError(f"failed to relax atoms at the start of the KMC iteration (system {w // self.G})")
                stats.relax_force_calls += int(st["force_calls"][w])
                s.basin = pos[w]
                s.e_basin32 = float(st["energy"][w])
                s.e_basin64 = float(st["energy64"][w])
                for m in s.walkers:
                    self._put(self._x0, m, s.basin)
                    mask[m] = 1
                s.phase = "search"
                start_search.extend(s.walkers)
            self.batch.set_basins(mask, self._x0)
        # searches that ended: collect valid events
        good = [w for w in searched if st["status"][w] == 0 and st["verified"][w] and st["new_basin"][w]]
        b2 = self.batch.walkers_positions(good) if good else {}
        sad = self.batch.walkers_saddles(good) if good else {}
        cap_hi = self.p.kmcMaxCap
        finished = []
        for w in searched:
            s = self._sys(w)
            r = st[w]
            stats.searches += 1
            stats.force_calls += int(r["force_calls"])
            stats.basin2_force_calls += int(r["basin2_force_calls"])
            if s.phase != "search":
                continue
            if r["status"] == 0:
                stats.converged += 1
                b32 = r["energy32"] - s.e_basin32
                b64 = r["energy64"] - s.e_basin64
                verified = bool(r["verified"]) and r["fmax64"] < self.fmax_verify and \
                    abs(b32 - b64) <= self.barrier_tol and r["lambda1"] < 0.0 < r["lambda2"] and b64 > 0.0
                if verified:
                    stats.verified += 1
                    if r["new_basin"]:
                        stats.new_basin += 1
                        if self.p.kmcMinCap <= b64 <= cap_hi:
                            rate = self._L.akmc_kmc_rate(float(b64), self.p.kmcPreFactor, self.p.kmcTemperature)
                            s.events.append(Event(s.e_basin64, float(r["energy_basin2_64"]), float(r["energy64"]),
                                                  float(b64), rate, np.array(s.site[w], copy=True), b2[w], sad[w],
                                                  w, int(r["force_calls"]), int(r["basin2_force_calls"])))
                            stats.valid_events += 1
                        else:
                            stats.out_of_window += 1
            if len(s.events) >= self.p.kmcImages:
                finished.append(s)
            else:
                start_search.append(w)   # the parent restarts the idle group (:425-437)
        # systems with enough events: stop their searches, take the KMC step
        for s in finished:
            if s.phase != "search":
                continue
            s.phase = "step"
            live = [m for m in s.walkers if m in self.pending]
            if live:
                self.batch.cancel(live)
                for m in live:
                    self.pending.discard(m)
            self._kmc_step(s, stats)
            if s.iteration >= self.p.kmcSteps:
                s.phase = "done"
            else:
                restart_relax.append(s)
        start_search = [w for w in start_search if self._sys(w).phase == "search" and w not in self.pending]
        if restart_relax:
            self._start_relaxes(restart_relax)
        self._start_searches(sorted(set(start_search)))

    # ------------------------------------------------------------ driver
    def run(self, max_micro_steps=10**9):
        """Run every system to kmcSteps KMC steps (or until max_micro_steps)."""
        stats = KMCStats()
        t0 = time.perf_counter()
        self._start_relaxes([s for s in self.systems if s.phase == "relax"])
        steps = 0
        while self.pending and steps < max_micro_steps:
            steps += self.batch.run(self.sync_every, sync_every=self.sync_every)
            self._harvest(stats)
        stats.micro_steps = steps
        stats.seconds = time.perf_counter() - t0
        return stats
