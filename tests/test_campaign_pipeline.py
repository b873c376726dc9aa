"""Host logic of the pipelined campaign loop, on the CPU: a scripted batch
plays the device (each walker's k-th search ends after a fixed number of
micro-steps with a result fixed by (walker, k)) and enforces the run-chunk
rules of include/akmc.h: one chunk in flight at a time, starts made while a
chunk is in flight take effect after it, statuses read during a chunk are
those of the last wait, and positions read during a chunk come from the
harvest of the chunk that was waited for.  The pipelined loop must record,
per walker, the same searches as the plain loop, and keep every walker busy.
"""
import numpy as np
import pytest

from akmc_amd import search as S
from akmc_amd import workloads as W

NW, NA = 6, 4  # walkers, atoms per walker
FIELDS = ["done", "status", "verified", "force_calls", "energy32", "energy64", "curvature32", "lambda1",
          "lambda2", "fmax64"]
DT = np.dtype([(f, np.float64) for f in FIELDS])


def length(w, k):
    """micro-steps of walker w's k-th search"""
    return 20 + 7 * ((3 * w + 5 * k) % 9)


class ScriptedBatch:
    def __init__(self):
        self.A = NW * NA
        self.offsets = np.arange(NW + 1) * NA
        self.natoms = np.full(NW, NA)
        self.inflight = False
        self.chunk = 0
        self.harvest_on = False
        self.searches = np.zeros(NW, dtype=int)  # searches started per walker
        self.left = np.zeros(NW, dtype=int)      # micro-steps left (0: not running)
        self.st = np.zeros(NW, dtype=DT)
        self.st["done"] = 1
        self.pos = np.zeros((NW, 3, NA))
        self.queued = []
        self.harvest = {}
        self.mirror = self.st.copy()

    def _start(self, w, x0):
        k = self.searches[w]
        self.searches[w] += 1
        self.left[w] = length(w, k)
        self.pos[w] = x0
        self.st[w] = 0
        self.st[w]["energy32"] = self.st[w]["energy64"] = 1.0 + 0.01 * k + 0.001 * w

    def start_flat(self, method, mask, x0, mode, frozen, params):
        for w in np.nonzero(mask)[0]:
            sl = slice(self.offsets[w], self.offsets[w + 1])
            xw = np.stack([x0[c * self.A:(c + 1) * self.A][sl] for c in range(3)])
            if self.inflight:
                self.queued.append((w, xw))  # runs from the next chunk
                self.mirror[w] = 0           # the host's copy shows it started
            else:
                self._start(w, xw)
        if not self.inflight:
            self.mirror = self.st.copy()

    def set_harvest(self, on=True):
        assert not self.inflight
        self.harvest_on = on

    def run_next(self, k):
        assert not self.inflight, "a run chunk is already in flight"
        self.inflight, self.chunk = True, k

    def run_async(self, k):
        self.run_next(k)

    def run_wait(self):
        assert self.inflight, "no run chunk in flight"
        ended = []
        for w in range(NW):
            if self.left[w] > 0:
                self.left[w] = max(0, self.left[w] - self.chunk)
                if self.left[w] == 0:
                    k = self.searches[w] - 1
                    r = self.st[w]
                    r["done"], r["status"], r["verified"], r["force_calls"] = 1, 0, 1, 100 + k
                    r["curvature32"], r["lambda1"], r["lambda2"], r["fmax64"] = -1.0, -1.0, 1.0, 0.01
                    self.pos[w, 0, 0] += 0.5 + 0.01 * k  # the saddle moves atom 0
                    ended.append(w)
        self.harvest = {w: self.pos[w].copy() for w in ended} if self.harvest_on else {}
        self.inflight = False
        for w, xw in self.queued:
            self._start(w, xw)
        self.queued = []
        self.mirror = self.st.copy()
        return self.chunk, int((self.left > 0).sum())

    def status_array(self):
        return self.mirror.copy()  # during a chunk: the last wait's states and the starts since

    def walkers_positions(self, ws):
        assert not self.inflight, "positions of a running batch"
        return {int(w): self.pos[w].copy() for w in ws}

    def harvested_positions(self, ws):
        for w in ws:
            assert int(w) in self.harvest, "not harvested by the last chunk"
        return {int(w): self.harvest[int(w)].copy() for w in ws}


def proposer(camp, w):
    basin = camp.basins[w]
    return basin + 0.05, np.ones((3, NA)), np.zeros(NA, np.int32), np.zeros(3)


def campaign():
    c = S.SaddleCampaign.__new__(S.SaddleCampaign)
    c.batch = ScriptedBatch()
    c.W = NW
    lat = np.diag([10.0, 10.0, 10.0]).ravel()
    c.cells = [W.Cell(np.zeros((3, NA)), np.zeros(NA, np.int32), lat) for _ in range(NW)]
    c.basins = [np.zeros((3, NA)) for _ in range(NW)]
    c.basin32 = np.ones(NW)
    c.basin64 = np.ones(NW)
    c.basin_ok = np.ones(NW, dtype=bool)
    c.method, c.params = "dimer", None
    c.fmax_verify, c.barrier_tol = 0.05, 0.01
    c.proposer = proposer
    c.found = [[] for _ in range(NW)]
    c.searches_per_basin, c.pool = None, None
    c.pool_idx = np.zeros(NW, dtype=np.int64)
    c._starts = np.zeros(NW, dtype=np.int64)
    c._catalogues, c._ctx, c._regions, c._next = {}, {}, {}, {}
    c.basin_moves = 0
    c._inflight = False
    c._harvest = False
    c.pending, c.cur_site = set(), {}
    c.off = c.batch.offsets
    c._x0 = np.zeros(3 * c.batch.A)
    c._mode = np.zeros(3 * c.batch.A)
    c._frozen = np.zeros(c.batch.A, dtype=np.int32)
    return c


def by_walker(records):
    out = {}
    for r in records:
        out.setdefault(r.walker, []).append((r.energy32, r.barrier64, tuple(r.atoms), r.disp.tobytes(), r.distinct))
    return out


def test_pipelined_loop_records_the_plain_loops_searches():
    plain = campaign()
    plain.begin()
    a = plain.advance(1000, sync_every=10)
    pipe = campaign()
    pipe.begin()
    b = S.CampaignStats()
    for _ in range(4):
        b.add(pipe.advance(250, sync_every=10, pipelined=True))
    assert pipe._inflight and pipe.batch.inflight
    pipe.drain(b, refill=False)
    assert not pipe.batch.inflight
    ra, rb = by_walker(a.records), by_walker(b.records)
    for w in range(NW):
        n = min(len(ra[w]), len(rb[w]))
        assert n >= 5
        assert ra[w][:n] == rb[w][:n], w
    # every search ran to its scripted length and each refill waited at most
    # the chunk in flight plus the rest of its own chunk
    for w in range(NW):
        done = len(rb[w])
        busy = sum(length(w, k) for k in range(done))
        assert busy <= 1000 + 10
        assert busy + done * 20 >= 1000 - max(length(w, k) for k in range(done + 1)) - 20
    # the plain loop can take over after a drain
    st = pipe.advance(50, sync_every=10)
    assert st.micro_steps == 50


def test_plain_loop_refuses_a_chunk_in_flight():
    c = campaign()
    c.begin()
    c.advance(20, sync_every=10, pipelined=True)
    with pytest.raises(RuntimeError):
        c.advance(20, sync_every=10)
    with pytest.raises(RuntimeError):
        c.harvest(S.CampaignStats())
    c.drain()
    c.advance(20, sync_every=10)
