"""Batched saddle searches with FP64 verification on the device.

``SaddleCampaign`` keeps W walker cells resident on the GPU, relaxes each
to its basin (FIRE, optim/Relax.cpp:44-315), then runs saddle searches
(dimer, or ART as optim/ART.cpp) from seeded local displacements around a
vacancy with atoms beyond ``region_radius`` frozen (the local active region
of optim/SaddleSearch.cpp:73-135, regionStyle 2).  Every walker is
refilled with a fresh search as soon as its previous one ends -- the role
of the reference's MPI task farm (kmc/KineticMonteCarlo.cpp:352-598).

Each converged search is re-verified inside the same device walker stream
in FP64: FP64 forces and energy at the saddle, then a Lanczos run on FP64
finite-difference Hessian products started on the search's mode.  A saddle
counts as FP64-verified when its largest FP64 atomic force is below
``fmax_verify``, its FP64 barrier is within ``barrier_tol`` (0.01 eV) of
the search's own barrier, the two lowest Lanczos curvatures bracket zero
(exactly one negative direction) and the barrier is positive.  Verified
saddles are de-duplicated per basin (barrier within 2 meV and the same
moved atoms within 0.15 A).

Basin turnover (``searches_per_basin`` = k): each walker owns a pool of
pre-relaxed basins -- the same cell shape and vacancy count, vacancies and
jitter re-drawn from seeds -- and moves to the next one after k searches,
as the reference's KMC loop moves on once a step's catalogue holds
kmcevents events (docs/examples/gdc/control: 5).  Each basin keeps its own
catalogue, so the distinct-saddle rate does not decay as a campaign runs
on (until the pool wraps around).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import lib as L
from . import workloads as W


def active_region(cell: W.Cell, pos, site, radius):
    """frozen mask: atoms farther than radius from site (minimum image)."""
    d = W.min_image(pos - np.asarray(site)[:, None], cell.lat)
    return (np.sqrt((d * d).sum(0)) >= radius).astype(np.int32)


@dataclass
class SaddleRecord:
    walker: int
    energy32: float
    energy64: float
    barrier32: float
    barrier64: float
    curvature: float
    lanczos: tuple
    fmax64: float
    atoms: np.ndarray
    disp: np.ndarray
    verified: bool = False
    distinct: bool = False
    site: np.ndarray | None = None  # vacancy site of the proposal
    active: np.ndarray | None = None      # free atoms of the search
    active_disp: np.ndarray | None = None  # their saddle displacement (3, n_active)
    pool: int = 0                         # index of the walker's basin in its pool


@dataclass
class CampaignStats:
    searches: int = 0
    converged: int = 0
    verified: int = 0
    distinct: int = 0
    force_calls: int = 0
    micro_steps: int = 0
    search_time: float = 0.0
    device_time: float = 0.0   # inside akmc_batch_run (device work + its host checks)
    harvest_time: float = 0.0  # reading ended searches
    refill_time: float = 0.0   # proposals and restarts
    # the parts of a chunk: launch, host work overlapped with the device
    # (record processing, proposal draws), waiting, reading ended searches,
    # restarting walkers
    t_launch: float = 0.0
    t_overlap: float = 0.0
    t_wait: float = 0.0
    t_collect: float = 0.0
    t_start: float = 0.0
    max_barrier_err: float = 0.0
    records: list = field(default_factory=list)
    # converged searches that failed one verification test (a search can fail several)
    fail_unchecked: int = 0   # the device verification did not complete
    fail_fmax: int = 0        # max FP64 force >= fmax_verify
    fail_barrier: int = 0     # |FP32 barrier - FP64 barrier| > barrier_tol
    fail_lambda1: int = 0     # lowest curvature not negative
    fail_lambda2: int = 0     # second curvature not positive
    fail_b64: int = 0         # FP64 barrier not positive

    _COUNTS = ("searches", "converged", "verified", "distinct", "force_calls", "micro_steps", "fail_unchecked",
               "fail_fmax", "fail_barrier", "fail_lambda1", "fail_lambda2", "fail_b64")

    def failures(self):
        return {k[5:]: getattr(self, k) for k in self._COUNTS if k.startswith("fail_")}

    def add(self, o: "CampaignStats"):
        for k in self._COUNTS:
            setattr(self, k, getattr(self, k) + getattr(o, k))
        self.search_time += o.search_time
        self.device_time += o.device_time
        for k in ("t_launch", "t_overlap", "t_wait", "t_collect", "t_start"):
            setattr(self, k, getattr(self, k) + getattr(o, k))
        self.harvest_time += o.harvest_time
        self.refill_time += o.refill_time
        self.max_barrier_err = max(self.max_barrier_err, o.max_barrier_err)
        self.records.extend(o.records)


class SaddleCampaign:
    def __init__(self, potentials_text, cells, method="dimer", region_radius=6.0, d0=0.1, params=None,
                 relax=None, seed=0, fmax_verify=0.05, barrier_tol=0.01, verify_nl=16, verify_delta=0.01,
                 proposer=None, skin=None, searches_per_basin=None):
        self.text = potentials_text
        self.cells = cells
        self.W = len(cells)
        self.method = method
        self.field = L.Field(potentials_text)
        if skin is not None:
            self.field.set_skin(skin)  # Verlet skin of the device lists (A)
        self.batch = L.Batch(self.field, cells)
        self.region_radius = region_radius
        self.d0 = d0
        self.params = params or (L.dimer_defaults() if method == "dimer" else L.art_defaults())
        self.relax = relax or L.relax_defaults()
        self.rng = np.random.default_rng(seed)
        self.fmax_verify = fmax_verify
        self.barrier_tol = barrier_tol
        self.verify_nl = verify_nl
        self.verify_delta = verify_delta
        self.proposer = proposer  # optional callable(campaign, w) -> (x0, mode, frozen, site)
        self.found = [[] for _ in range(self.W)]
        # basin turnover: pool[w] = [(cell, basin, basin32, basin64, ok)], the
        # pool index of the walker's current basin, searches started on it,
        # and the catalogues of the basins it left (by pool index)
        self.searches_per_basin = searches_per_basin
        self.pool = None
        self.pool_idx = np.zeros(self.W, dtype=np.int64)
        self._starts = np.zeros(self.W, dtype=np.int64)
        self._catalogues = {}
        self._ctx = {}  # walker -> (pool index, basin32, basin64, basin) of its running search
        self.basin_moves = 0
        self._regions = {}
        self._next = {}  # walker -> its next proposal, drawn ahead
        self._inflight = False  # a pipelined chunk is running (advance(pipelined=True))
        self._harvest = False  # chunks copy the ended walkers' positions to host memory
        self.pending = set()
        self.cur_site = {}
        A = self.batch.A
        self.off = self.batch.offsets
        self._x0 = np.zeros(3 * A)
        self._mode = np.zeros(3 * A)
        self._frozen = np.zeros(A, dtype=np.int32)

    # ---------------------------------------------------------- basins
    def relax_basins(self, sync_every=50, max_steps=5000):
        self.batch.set_frozen([np.zeros(c.n, np.int32) for c in self.cells])
        self.batch.start_relax(self.relax)
        self.batch.run(max_steps, sync_every=sync_every)
        st = self.batch.status()
        self.basin_ok = np.array([s["status"] == 0 for s in st])
        self.basins = self.batch.positions()
        self._regions = {}
        self._next = {}
        self.batch.save_basin()
        self.basin32 = self.batch.search_energies(32)
        self.basin64 = self.batch.search_energies(64)
        self.batch.set_verify(self.verify_nl, self.verify_delta)
        return self.basin_ok

    def relax_pool(self, size, seed, sync_every=50, max_steps=5000):
        """Relax `size` basins per walker (pool entry 0 is the campaign's own
        cells, as relax_basins leaves them; entries 1.. re-draw the vacancy
        sites and jitter of the same cell shape from seeds) for basin
        turnover.  Leaves the batch on pool entry 0."""
        first = (list(self.cells), [b.copy() for b in self.basins], self.basin32.copy(), self.basin64.copy(),
                 self.basin_ok.copy())
        rng = np.random.default_rng(seed)
        rounds = [first]
        for _ in range(1, size):
            cells = []
            for c in first[0]:
                m = c.meta
                cells.append(W.sc_cell(m["nx"], m["ny"], m["nz"], n_vac=len(c.vacancies), jitter=m.get("jitter", 0.01),
                                       seed=int(rng.integers(0, 2**31 - 1))))
            self.cells = cells
            self.batch.upload(cells)
            self.relax_basins(sync_every=sync_every, max_steps=max_steps)
            rounds.append((cells, [b.copy() for b in self.basins], self.basin32.copy(), self.basin64.copy(),
                           self.basin_ok.copy()))
        self.pool = [[(r[0][w], r[1][w], float(r[2][w]), float(r[3][w]), bool(r[4][w])) for r in rounds]
                     for w in range(self.W)]
        self.pool_idx[:] = 0
        self._starts[:] = 0
        self._use_pool_entry(range(self.W), upload=True)

    def _use_pool_entry(self, walkers, upload=False):
        """make each listed walker's current pool entry its basin (host
        state); with upload, put every walker on its basin on the device"""
        for w in walkers:
            c, basin, b32, b64, ok = self.pool[w][int(self.pool_idx[w]) % len(self.pool[w])]
            self.cells[w] = c
            self.basins[w] = basin
            self.basin32[w] = b32
            self.basin64[w] = b64
            self.basin_ok[w] = ok
        if upload:
            self.batch.upload([W.Cell(p, c.types, c.lat) for p, c in zip(self.basins, self.cells)])
            self.batch.save_basin()

    def _advance_basin(self, w):
        """after searches_per_basin searches the walker moves to its next
        pooled basin, with that basin's own catalogue"""
        k = self.searches_per_basin
        if self.pool is None or not k or self._starts[w] < k:
            return
        self._catalogues[(w, int(self.pool_idx[w]) % len(self.pool[w]))] = self.found[w]
        self.pool_idx[w] += 1
        self._starts[w] = 0
        self._use_pool_entry([w])
        self.found[w] = self._catalogues.get((w, int(self.pool_idx[w]) % len(self.pool[w])), [])
        self._regions = {key: v for key, v in self._regions.items() if key[0] != w}
        self.basin_moves += 1

    def _catalogue(self, w, pool):
        """the de-duplication list of pool entry `pool` of walker w"""
        if pool == int(self.pool_idx[w]) % (len(self.pool[w]) if self.pool else 1):
            return self.found[w]
        return self._catalogues.setdefault((w, pool), [])

    # ---------------------------------------------------------- searches
    def _site_region(self, w, v):
        """(frozen mask, the 12 neighbours) of vacancy v of walker w at its
        basin; fixed per basin, so computed once."""
        key = (w, int(self.pool_idx[w]), v)
        hit = self._regions.get(key)
        if hit is None:
            c = self.cells[w]
            site = c.vacancies[v]
            r = W.site_distances(W.Cell(self.basins[w], c.types, c.lat), site)  # both from one pass
            frozen = (r >= self.region_radius).astype(np.int32)
            nb = W.neighbours_of_site(c, site, 12, r=r)
            hit = self._regions[key] = (frozen, nb)
        return hit

    def proposal(self, w):
        """Seeded random displacement (d0, centre-of-mass free) of the 12
        neighbours of one of the walker's vacancies; the displacement
        direction is the start mode."""
        c = self.cells[w]
        v = int(self.rng.integers(0, len(c.vacancies)))
        frozen, nb = self._site_region(w, v)
        basin = self.basins[w]
        disp, mode = W.local_displacement(W.Cell(basin, c.types, c.lat), nb, self.d0,
                                          int(self.rng.integers(0, 2**31 - 1)))
        return basin + disp, mode, frozen, c.vacancies[v]

    def _draw(self, w):
        """the walker's next proposal, on the basin its next search runs on
        (moving to the next pooled basin first when the current one has had
        its searches); returns (x0, mode, frozen, site, context)"""
        self._advance_basin(w)
        pool = int(self.pool_idx[w]) % (len(self.pool[w]) if self.pool else 1)
        ctx = (pool, self.basin32[w], self.basin64[w], self.basins[w], self.cells[w].lat)
        self._starts[w] += 1
        x0, mode, frozen, site = (self.proposer or SaddleCampaign.proposal)(self, w)
        return x0, mode, frozen, site, ctx

    def precompute(self, walkers):
        """draw the next proposal of each listed walker now (random proposals
        do not depend on how the current search ends), so a refill between
        device chunks only uploads arrays"""
        if self.proposer is not None:
            return
        for w in walkers:
            if w not in self._next and self.basin_ok[w]:
                self._next[w] = self._draw(w)

    def start(self, walkers):
        walkers = [int(w) for w in walkers if self.basin_ok[w]]
        if not walkers:
            return
        A = self.batch.A
        mask = np.zeros(self.W, dtype=np.int32)
        for w in walkers:
            nxt = self._next.pop(w, None) if self.proposer is None else None
            x0, mode, frozen, site, ctx = nxt if nxt is not None else self._draw(w)
            self.cur_site[w] = site
            self._ctx[w] = ctx
            a, b = self.off[w], self.off[w + 1]
            for c in range(3):
                self._x0[c * A + a:c * A + b] = x0[c]
                self._mode[c * A + a:c * A + b] = mode[c]
            self._frozen[a:b] = frozen
            mask[w] = 1
            self.pending.add(w)
        self.batch.start_flat(self.method, mask, self._x0, self._mode, self._frozen, self.params)

    def harvest(self, stats: CampaignStats):
        """Collect ended searches (their FP64 verification already ran on
        the device); returns the walkers that ended."""
        if self._inflight:
            raise RuntimeError("a pipelined chunk is in flight: drain() first")
        done, raw = self.collect(stats)
        self.process(raw, stats)
        return done

    def collect(self, stats: CampaignStats):
        """The device part of a harvest: statuses and saddle positions of
        the ended searches, with the context their records need (site, free
        atoms), taken before the walkers are refilled.  Returns (ended
        walkers, raw records for process())."""
        st = self.batch.status_array()
        done = [w for w in sorted(self.pending) if st["done"][w]]
        good = [w for w in done if st["status"][w] == 0 and st["verified"][w]]
        if not good:
            positions = {}
        elif self._inflight:  # copied to host memory at the end of the chunk that was waited for
            positions = self.batch.harvested_positions(good)
        else:
            positions = self.batch.walkers_positions(good)
        raw = []
        for w in done:
            s = st[w]
            self.pending.discard(w)
            stats.searches += 1
            stats.force_calls += int(s["force_calls"])
            if s["status"] != 0:
                continue
            stats.converged += 1
            if not s["verified"]:
                stats.fail_unchecked += 1
                continue
            act = np.nonzero(self._frozen[self.off[w]:self.off[w + 1]] == 0)[0]
            raw.append((w, s, positions[w], self.cur_site.get(w), act, self._ctx.get(w)))
        return done, raw

    def process(self, raw, stats: CampaignStats):
        """The host part of a harvest: FP64 tests, records, de-duplication
        (needs nothing from the device, so it can run while the next chunk
        does)."""
        for (w, s, pos, site, act, ctx) in raw:
            pool, e32, e64, basin, lat = ctx
            b32 = s["energy32"] - e32
            b64 = s["energy64"] - e64
            err = abs(b32 - b64)
            stats.max_barrier_err = max(stats.max_barrier_err, err)
            d = W.min_image(pos - basin, lat)
            dn = np.sqrt((d * d).sum(0))
            atoms = np.nonzero(dn > 0.3)[0]
            rec = SaddleRecord(w, s["energy32"], s["energy64"], float(b32), float(b64), s["curvature32"],
                               (s["lambda1"], s["lambda2"]), s["fmax64"], atoms, d[:, atoms],
                               site=site, active=act, active_disp=d[:, act], pool=pool)
            tests = dict(fmax=s["fmax64"] < self.fmax_verify, barrier=err <= self.barrier_tol,
                         lambda1=s["lambda1"] < 0.0, lambda2=s["lambda2"] > 0.0, b64=b64 > 0.0)
            for k, ok in tests.items():
                if not ok:
                    setattr(stats, "fail_" + k, getattr(stats, "fail_" + k) + 1)
            rec.verified = all(tests.values())
            if rec.verified:
                stats.verified += 1
                cat = self._catalogue(w, pool)
                rec.distinct = all(not self._same(rec, o) for o in cat)
                if rec.distinct:
                    cat.append(rec)
                    stats.distinct += 1
            stats.records.append(rec)

    @staticmethod
    def _same(a: SaddleRecord, b: SaddleRecord):
        if abs(a.barrier64 - b.barrier64) > 2e-3 or a.atoms.size != b.atoms.size:
            return False
        if not np.array_equal(a.atoms, b.atoms):
            return False
        return float(np.abs(a.disp - b.disp).max()) < 0.15

    def begin(self):
        """Start a search on every walker."""
        self.start(range(self.W))

    def advance(self, micro_steps, sync_every=25, refill=True, pipelined=False):
        """Run micro_steps micro-steps (a multiple of sync_every), refilling
        walkers whose search ended; returns the searches that ended in this
        window.  Searches still running carry over to the next window.

        Each chunk of sync_every micro-steps runs on the device while the
        host processes the previous chunk's records and draws the next
        proposals; between chunks the host only reads the ended searches
        and uploads the drawn proposals.  Pipelined, the next chunk is
        launched before that: the device does not wait for the host, and a
        search that ends waits one more chunk for its refill (a chunk stays
        in flight between calls; drain() ends it)."""
        if pipelined:
            return self._advance_pipelined(micro_steps, sync_every, refill)
        if self._inflight:
            raise RuntimeError("a pipelined chunk is in flight: drain() first")
        stats = CampaignStats()
        t0 = time.perf_counter()
        steps = 0
        raw = []
        while steps < micro_steps and self.pending:
            k = min(sync_every, micro_steps - steps)
            t1 = time.perf_counter()
            self.batch.run_async(k)
            ta = time.perf_counter()
            self.process(raw, stats)
            tb = time.perf_counter()
            if refill:
                self.precompute(sorted(self.pending))
            tc = time.perf_counter()
            got, _ = self.batch.run_wait()
            steps += got
            t2 = time.perf_counter()
            done, raw = self.collect(stats)
            t3 = time.perf_counter()
            if refill and done:
                self.start(done)
            t4 = time.perf_counter()
            stats.device_time += (t2 - tc) + (ta - t1)
            stats.harvest_time += (t3 - t2) + (tb - ta)
            stats.refill_time += (t4 - t3) + (tc - tb)
            stats.t_launch += ta - t1
            stats.t_overlap += tc - ta
            stats.t_wait += t2 - tc
            stats.t_collect += t3 - t2
            stats.t_start += t4 - t3
            if got == 0:
                break
        self.process(raw, stats)
        stats.search_time = time.perf_counter() - t0
        stats.micro_steps = steps
        return stats

    def _advance_pipelined(self, micro_steps, sync_every, refill):
        stats = CampaignStats()
        t0 = time.perf_counter()
        steps = 0
        raw = []
        if not self._harvest:
            self.batch.set_harvest(True)
            self._harvest = True
        if not self._inflight and self.pending:
            self.batch.run_next(sync_every)
            self._inflight = True
        while steps < micro_steps and self._inflight:
            ta = time.perf_counter()
            self.process(raw, stats)
            tb = time.perf_counter()
            if refill:
                self.precompute(sorted(self.pending))
            tc = time.perf_counter()
            got, _ = self.batch.run_wait()
            self._inflight = False
            steps += got
            t1 = time.perf_counter()
            # the next chunk at once; the walkers that ended wait through it
            if self.pending:
                self.batch.run_next(sync_every)
                self._inflight = True
            t2 = time.perf_counter()
            done, raw = self.collect(stats)
            t3 = time.perf_counter()
            if refill and done:
                self.start(done)
                if not self._inflight:
                    self.batch.run_next(sync_every)
                    self._inflight = True
            t4 = time.perf_counter()
            stats.device_time += t1 - tc
            stats.harvest_time += (t3 - t2) + (tb - ta)
            stats.refill_time += (t4 - t3) + (tc - tb)
            stats.t_launch += t2 - t1
            stats.t_overlap += tc - ta
            stats.t_wait += t1 - tc
            stats.t_collect += t3 - t2
            stats.t_start += t4 - t3
            if got == 0 and not done:
                break
        self.process(raw, stats)  # beside the chunk left in flight
        stats.search_time = time.perf_counter() - t0
        stats.micro_steps = steps
        return stats

    def drain(self, stats: CampaignStats = None, refill=True):
        """End a pipelined run: wait for the chunk in flight, record the
        searches it ended and (refill) start their walkers' next ones."""
        stats = stats if stats is not None else CampaignStats()
        if self._inflight:
            self.batch.run_wait()
            self._inflight = False
            done, raw = self.collect(stats)
            if refill and done:
                self.start(done)
            self.process(raw, stats)
        return stats

    def run(self, sync_every=25, max_steps=10**9):
        """One search per walker, run until all have ended."""
        self.begin()
        stats = CampaignStats()
        t0 = time.perf_counter()
        steps = 0
        while self.pending and steps < max_steps:
            steps += self.batch.run(sync_every, sync_every=sync_every)
            self.harvest(stats)
        stats.search_time = time.perf_counter() - t0
        stats.micro_steps = steps
        return stats
