#!/usr/bin/env python3
"""Benchmark of the force-and-saddle-search path on one MI355X.

Workload (value): batched dimer searches on Sutton-Chen Cu walkers --
512 cells of 500-4000 atoms (FCC n_x x n_y x n_z, 1-3 vacancies at seeded
sites, N(0, 0.01 A) jitter), each relaxed to its basin (FIRE) before the
timed region.  Every walker runs dimer searches from seeded 0.1 A
displacements of a vacancy's 12 neighbours and is refilled with a new
search as soon as one ends; each converged saddle is re-verified in FP64
on the device (max force, FP64 barrier within 0.01 eV of the search's, one
negative Lanczos curvature) and de-duplicated per basin.  One step = one
window of 500 batched micro-steps (a micro-step is one force evaluation of
every running walker plus its state update).

value = FP64-verified distinct saddles ended in the timed windows / their
wall time.  Basins turn over: every walker owns a pool of pre-relaxed
basins (same cell shape and vacancy count, re-drawn vacancy sites and
jitter) and moves to the next after --searches-per-basin searches (5, the
kmcevents of the reference's example control file), each basin with its
own de-duplication catalogue, so the rate is a steady state rather than a
catalogue that saturates as the run goes on.
Searches evaluate their active zones (the free atoms' forces, the
densities and energies within the list radius of a free atom), bit for bit
the results of full evaluations.
Also reported: the plain unbatched GPU dimer (one walker, host-driven
micro-steps) on the same cells, force evaluations per second, the roofline
of the dominant kernel of the search micro-steps (the zoned Sutton-Chen
pass 1) timed with HIP events on the batch's stream, with the full
evaluation's pass 2 beside it, the learned-proposal A/B test, the KMC step
(64 systems x 8 walkers taking kmcSteps steps of 5 valid events each), the
tabulated-EAM field's force evaluation on the same walkers, the ionic
(rock-salt Ewald) and full-loop (31,980-atom alloy) sections, and the CPU
baseline (the FP64 C restatement of the reference, OpenMP) on a bounded
sample of the same searches.

    python bench.py [--gpus 1] [--steps 3] [--warmup 2]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64-verified saddles/sec at ≤0.01 eV barrier error; force evals/sec vs CPU"
FP32_VECTOR_PEAK_TFLOPS = 157.3
FP64_VECTOR_PEAK_TFLOPS = 78.6
# algorithmic FLOP per ordered in-cutoff pair of the zoned pass 1 of the
# search micro-steps: the density's 20 (1/r^2, (a/r)^2, (a/r)^6, the
# coefficient, the sum, with the FP64 displacement and cutoff test) plus 5
# for the pair energy (a/r)^9 = (a/r)^8 sqrt((a/r)^2), 1/2 eps (a/r)^9, its sum
ZONE_PASS1_FLOP = 25
HBM_PEAK_GBS = 8000.0
# HBM bytes per ordered pair of one all-walker k_sc_force launch, from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on `bench.py --forces-only 5`,
# calibrated on a 1 GiB device copy (FETCH_SIZE reads half the copied bytes,
# WRITE_SIZE all of them); the summary is under profiles/
TRAFFIC_BYTES_PER_PAIR = 3.35
# the same for the zoned pass 1 of the search micro-steps (k_sc_density with
# pair energies), from the counter runs of `bench.py --zone-passes 10`
ZONE_TRAFFIC_BYTES_PER_PAIR = 9.72


def ordered_pairs(cells, positions, rc):
    """Ordered in-cutoff pairs (i, j != i, r <= rc) per walker, minimum image."""
    from scipy.spatial import cKDTree
    P = []
    for c, p in zip(cells, positions):
        L = np.array([c.lat[0], c.lat[4], c.lat[8]])
        q = np.mod(p.T, L)
        q = np.where(q >= L, 0.0, q)
        t = cKDTree(q, boxsize=L)
        P.append(int(t.count_neighbors(t, rc)) - c.n)
    return np.array(P)


def atom_pairs(cells, positions, rc):
    """Per walker, each atom's in-cutoff neighbours (r <= rc, minimum image)."""
    from scipy.spatial import cKDTree
    out = []
    for c, p in zip(cells, positions):
        L = np.array([c.lat[0], c.lat[4], c.lat[8]])
        q = np.mod(p.T, L)
        q = np.where(q >= L, 0.0, q)
        t = cKDTree(q, boxsize=L)
        out.append(np.asarray(t.query_ball_point(q, rc, return_length=True)) - 1)
    return out


def zone_work(camp, cells, rc):
    """The algorithmic work of one evaluation of the search walkers as the
    micro-steps run it: ordered in-cutoff pairs of pass 1 (each walker's
    zone, or all its atoms when it evaluates in full) and of pass 2 (its
    free atoms, or all), over the walkers in FP32 mode, at the current
    probes; zone atoms and free atoms per walker."""
    member, size, flags, active = camp.batch.zone()
    probes, _ = camp.batch.download()
    cnt = atom_pairs(cells, probes, rc)
    p1 = p2 = n_zone = n_free = n_w = 0
    for w, c in enumerate(cells):
        if active[w] != 1:
            continue
        o, k = camp.off[w], cnt[w]
        n_w += 1
        if (flags[w] & 1) and not (flags[w] & 2):
            zm = member[o:o + c.n] != 0
            fr = camp._frozen[o:o + c.n] == 0
            p1 += int(k[zm].sum())
            p2 += int(k[fr].sum())
            n_zone += int(zm.sum())
            n_free += int(fr.sum())
        else:
            p1 += int(k.sum())
            p2 += int(k.sum())
            n_zone += c.n
            n_free += c.n
    return dict(pairs_pass1=p1, pairs_pass2=p2, zone_atoms=n_zone, free_atoms=n_free, walkers=n_w)


def plain_dimer_rate(L, W, text, cells, basins, proposal_fn, params, n_search, basin32, basin64, eager=False):
    """The plain GPU dimer: one walker at a time, host-synchronised every
    micro-step (the reference's control flow with device forces), with the
    same FP64 verification."""
    f = L.Field(text)
    t_total, ok, calls = 0.0, 0, 0
    for k in range(n_search):
        w = k % len(cells)
        c = cells[w]
        b = L.Batch(f, [W.Cell(basins[w], c.types, c.lat)])
        if eager:
            b.set_graphs(False)
        b.forces()
        b.autotune(5)
        x0, mode, frozen, _ = proposal_fn(w)
        t0 = time.perf_counter()
        b.set_verify(16, 0.01)
        b.start_dimer([mode], params=params, x0_per_walker=[x0], frozen_per_walker=[frozen])
        b.run(10**6, sync_every=1)
        st = b.status()[0]
        t_total += time.perf_counter() - t0
        b32 = st["energy32"] - basin32[w]
        b64 = st["energy64"] - basin64[w]
        ok += (st["status"] == 0 and st["verified"] and st["fmax64"] < 0.05 and abs(b32 - b64) <= 0.01
               and st["lambda1"] < 0 < st["lambda2"])
        calls += st["force_calls"]
        b.close()
    return ok, t_total, calls


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _omp_threads(n):
    """set the OpenMP thread count of the oracle's fast build (libgomp)"""
    import ctypes
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def cpu_force_timing(cells, basins, cores, repeats=10, warmup=3):
    """median wall time of one FP64 Sutton-Chen force evaluation of the
    restatement (steady_clock-like perf_counter, `warmup` untimed calls, then
    the median of `repeats`) on the median-size walker cell, at `cores`
    OpenMP threads and at 1 thread"""
    from oracle import oracle as O
    from akmc_amd import workloads as W
    cu = W.SC_CU
    f = O.sc_field(["Cu"], [("Cu", "Cu", cu["eps"], cu["a"], cu["n"], cu["m"], cu["c"])], 2 * cu["a"])
    w = int(np.argsort([c.n for c in cells])[len(cells) // 2])
    c = cells[w]
    out = {"cell_atoms": c.n}
    for label, nt in (("all_cores", cores), ("one_core", 1)):
        _omp_threads(nt)
        ts = []
        for k in range(warmup + repeats):
            t0 = time.perf_counter()
            O.sc_forces(f, basins[w], c.types, c.lat, omp=True, fast=True)
            if k >= warmup:
                ts.append(time.perf_counter() - t0)
        out[label] = {"threads": nt, "median_ms": round(1e3 * float(np.median(ts)), 3),
                      "force_evals_per_s": round(1.0 / float(np.median(ts)), 2)}
    _omp_threads(cores)
    return out


def cpu_baseline(cells, basins, proposal_fn, budget_s=20.0, max_searches=3):
    """FP64 C restatement of the reference (OpenMP Sutton-Chen force, as the
    reference's metal build) running the same dimer searches."""
    from oracle import oracle as O
    from akmc_amd import workloads as W
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    _omp_threads(cores)
    cu = W.SC_CU
    f = O.sc_field(["Cu"], [("Cu", "Cu", cu["eps"], cu["a"], cu["n"], cu["m"], cu["c"])], 2 * cu["a"])
    order = np.argsort([c.n for c in cells])
    pick = [int(order[len(order) // 2 + k]) for k in range(max_searches)]
    t_total, ok, calls, done = 0.0, 0, 0, 0
    for w in pick:
        c = cells[w]
        h = O.FieldHandle(f, 2, c.types, c.lat, fast=True)
        x0, mode, frozen, _ = proposal_fn(w)
        m = mode[:, frozen == 0].T.reshape(-1)
        t0 = time.perf_counter()
        r = O.dimer_search(h, x0, frozen, m, O.dimer_defaults(fast=True), fast=True)
        t_total += time.perf_counter() - t0
        ok += r["status"] == 0
        calls += r["fcalls"]
        done += 1
        if t_total > budget_s:
            break
    sizes = [cells[w].n for w in pick[:done]]
    return dict(saddles_per_s=ok / t_total if t_total > 0 else 0.0, force_evals_per_s=calls / t_total,
                cores=cores, searches=done, converged=ok, seconds=t_total, sizes=sizes,
                force_timing=cpu_force_timing(cells, basins, cores), model=cpu_model())


def copy_calibration(mib):
    """A plain device-to-device copy of known size: reading FETCH_SIZE and
    WRITE_SIZE for it gives the scale from counter units to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    c = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        c.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": n * 4, "launches": 5}), flush=True)


def forces_only(L, W, args):
    """Force evaluations of every walker of the batched workload at its
    seeded (unrelaxed) positions, lanes per atom autotuned."""
    cu = W.SC_CU
    text = L.sc_potentials_text([("Cu", "Cu", cu["eps"], cu["a"], cu["n"], cu["m"], cu["c"])], 2 * cu["a"])
    cells = W.batched_sc_cells(n_walkers=args.walkers, seed=args.seed)
    b = L.Batch(L.Field(text), cells)
    b.forces()
    variant, _ = b.autotune(5)
    ph = b.time_phases(args.forces_only)
    P = int(ordered_pairs(cells, [c.pos for c in cells], 2 * cu["a"]).sum())
    print(json.dumps({"forces_only": args.forces_only, "walkers": len(cells), "atoms": b.A,
                      "pairs": P, "variant": variant, "phase_ms": [round(float(x), 4) for x in ph]}),
          flush=True)


def zone_passes(L, S, W, args):
    """The zoned pass kernels of the search micro-steps alone, for hardware
    counter runs: the batched campaign (3 pooled basins per walker, kernels
    launched one by one) after a 250-micro-step warm window and the per-pass
    variant tuning, then args.zone_passes evaluations of the running searches
    at their probes."""
    cu = W.SC_CU
    rc = 2 * cu["a"]
    text = L.sc_potentials_text([("Cu", "Cu", cu["eps"], cu["a"], cu["n"], cu["m"], cu["c"])], rc)
    cells = W.batched_sc_cells(n_walkers=args.walkers, seed=args.seed)
    camp = S.SaddleCampaign(text, cells, params=L.dimer_defaults(maxParIter=400), seed=args.seed)
    camp.batch.set_graphs(False)  # counter and trace tools do not follow graph launches
    camp.batch.forces()
    camp.batch.autotune(5)
    camp.relax_basins()
    camp.relax_pool(3, seed=args.seed + 17)
    camp.searches_per_basin = args.searches_per_basin
    camp.begin()
    camp.advance(250)
    sv, _ = camp.batch.autotune_search(3)
    camp.batch.time_search_phases(1)
    zw = zone_work(camp, cells, rc)
    ph = camp.batch.time_search_phases(args.zone_passes)
    print(json.dumps({"zone_passes": args.zone_passes, "search_variant": sv, "zones": zw,
                      "phase_ms": [round(float(x), 4) for x in ph]}), flush=True)


def drain(camp, stats):
    """Let every running search end (no refill)."""
    while camp.pending:
        camp.harvest(stats)
        if camp.pending:
            camp.batch.run(25, sync_every=25)


def learned_ab(camp, records, window, epochs, train_saddles):
    """Config 4 on the bench's own walkers: train the MLP on FP64-verified
    saddles of the batched campaign (the timed windows' and, until
    train_saddles are in, further windows of the same campaign), time its
    MFMA kernels, then run the same cells from random and from learned
    proposals (fresh de-duplication, one warm-up and one timed window
    each)."""
    from akmc_amd import learn as LR
    from akmc_amd import search as S
    recs = [r for r in records if r.verified]
    t0 = time.perf_counter()
    extra = 0
    while len(recs) < train_saddles:
        st = camp.advance(window, sync_every=25)
        recs.extend(r for r in st.records if r.verified)
        extra += 1
    t_gen = time.perf_counter() - t0
    drain(camp, S.CampaignStats())
    X, T = LR.dataset_from_records(camp, recs)
    net = LR.MLP(max_rows=8192, seed=1)
    net.fit_scaling(X)
    t0 = time.perf_counter()
    hist = net.train(X, T, epochs=epochs, batch=4096, lr=2e-3, seed=3)
    t_train = time.perf_counter() - t0
    rows = 8192
    fwd_ms, fb_ms = net.time(rows, 20)
    fwd_flop = 2.0 * 22912 * rows               # 48x128 + 128x128 + 128x3 MACs per row
    fb_flop = fwd_flop + 2.0 * 39680 * rows     # + dZ2, dZ1 and the three weight gradients
    out = {"training_rows": int(X.shape[0]), "training_saddles": len(recs),
           "training_data": f"FP64-verified saddles of the batched campaign ({extra} windows beyond the timed ones, "
                            f"{t_gen:.1f} s)",
           "epochs": epochs, "train_seconds": round(t_train, 2),
           "loss_first_last": [round(float(hist[0]), 5), round(float(hist[-1]), 5)],
           "mlp_mfma": {"rows": rows, "forward_ms": round(fwd_ms, 4), "forward_backward_ms": round(fb_ms, 4),
                        "forward_tflops": round(fwd_flop / (fwd_ms * 1e-3) / 1e12, 2),
                        "forward_backward_tflops": round(fb_flop / (fb_ms * 1e-3) / 1e12, 2),
                        "peak_bf16_dense_tflops": 2500.0}}
    k_turnover = camp.searches_per_basin
    camp.searches_per_basin = None  # both arms search the same basins
    for name, prop in (("random", None), ("learned", LR.LearnedProposer(camp, net, seed=4))):
        drain(camp, S.CampaignStats())
        camp.proposer = prop
        camp.found = [[] for _ in range(camp.W)]
        camp.begin()
        camp.advance(window // 2, sync_every=25)
        st = camp.advance(window, sync_every=25)
        out[name] = {"verified_saddles_per_s": round(st.verified / st.search_time, 2),
                     "distinct_saddles_per_s": round(st.distinct / st.search_time, 2),
                     "searches": st.searches, "verified": st.verified, "distinct": st.distinct,
                     "force_calls_per_verified": round(st.force_calls / max(st.verified, 1), 1)}
        print(f"[bench] {name} proposals: {out[name]}", file=sys.stderr, flush=True)
    camp.proposer = None
    camp.searches_per_basin = k_turnover
    return out


def eam_section(L, W, cells, basins, P, args):
    """The tabulated-EAM field (LAMMPS setfl tables sampling the Cu
    Sutton-Chen functions on a 1e-3 mesh) on the same 512 walkers at their
    basins: one force evaluation of every walker, per-pass device times
    (HIP events) and ordered pairs per second of the faster of the unstaged
    and LDS-staged kernels (both totals reported)."""
    import tempfile
    cu = W.SC_CU
    rc = 2 * cu["a"]
    path = os.path.join(tempfile.mkdtemp(), "cu_sc.eam")
    W.write_sc_setfl(path, cu["eps"], cu["a"], cu["n"], cu["m"], cu["c"], rc)
    f = L.Field(L.eam_potentials_text([("Cu", "Cu", path)], rc))
    b = L.Batch(f, [W.Cell(p, c.types, c.lat) for p, c in zip(basins, cells)])
    b.forces()
    ph_global = b.time_phases(20)
    b.set_staging(True)  # the walkers' gather records in LDS (the tables stay in global memory)
    b.forces()
    ph_staged = b.time_phases(20)
    staged = bool(ph_staged.sum() < ph_global.sum())
    ph = ph_staged if staged else ph_global
    Ptot = int(P.sum())
    out = {"table": "LAMMPS setfl, Cu Sutton-Chen functions sampled at dr = drho = 1e-3",
           "lds_staged": staged, "force_eval_ms_all_walkers_unstaged": round(float(ph_global.sum()), 4),
           "force_eval_ms_all_walkers_staged": round(float(ph_staged.sum()), 4),
           "ms_lists": round(float(ph[0]), 4), "ms_density": round(float(ph[1]), 4),
           "ms_force": round(float(ph[2]), 4), "ms_reduce": round(float(ph[3]), 4),
           "force_eval_ms_all_walkers": round(float(ph.sum()), 4),
           "ordered_pairs": Ptot,
           "pairs_per_s_force_pass": round(Ptot / (ph[2] * 1e-3), 1),
           "force_evals_per_s_cells_device": round(len(cells) / (float(ph.sum()) * 1e-3), 1)}
    b.close()
    f.close()
    return out


def kmc_section(L, W, text, args):
    """The KMC step around the search path: 64 independent systems (walker
    cells of the batched workload) x 8 walkers each, kmcSteps 3 of
    kmcevents 5 (docs/examples/gdc/control's count) per system.  Each
    search is FP64-verified and followed on the device by its basin-2 kick,
    FIRE relax and checkBasin; each system takes its KMC step (rates,
    choice, clock, move) as soon as it holds 5 valid events, then re-relaxes
    its new basin 1 and searches again."""
    from akmc_amd import kmc as M
    systems = W.batched_sc_cells(n_walkers=args.kmc_systems, seed=args.seed + 1)
    p = M.KMCParams(kmcSteps=args.kmc_steps, kmcImages=5)
    run = M.KineticMonteCarlo(text, systems, walkers_per_system=8, params=p,
                              search_params=L.dimer_defaults(maxParIter=400), seed=args.seed, skin=args.skin)
    if args.eager:
        run.batch.set_graphs(False)
    st = run.run()
    ea = [r.activation_energy for r in st.steps]
    out = {"systems": len(systems), "walkers": run.W, "atoms": run.A, "kmc_steps": st.kmc_steps,
           "seconds": round(st.seconds, 2), "kmc_steps_per_s": round(st.kmc_steps / st.seconds, 2),
           "valid_events_per_s": round(st.valid_events / st.seconds, 1),
           "searches": st.searches, "converged": st.converged, "verified": st.verified,
           "new_basin": st.new_basin, "valid_events": st.valid_events, "out_of_window": st.out_of_window,
           "force_calls_per_valid_event": round(st.force_calls / max(st.valid_events, 1), 1),
           "basin2_force_calls_per_new_basin": round(st.basin2_force_calls / max(st.new_basin, 1), 1),
           "basin1_relax_force_calls": st.relax_force_calls, "micro_steps": st.micro_steps,
           "mean_chosen_barrier_eV": round(float(np.mean(ea)), 4) if ea else None,
           "simulated_time_s_mean": float(np.mean([s.time for s in run.systems]))}
    run.batch.close()
    return out


def ion_pairs_with_images(c, rc):
    """unordered in-cutoff pairs of a cell including periodic images (the
    reference's no-image loop over a <= b and image shifts): half the
    ordered (i, j-image) pairs plus the self-image pairs, counted over the
    27 nearest image cells (rc below the shortest cell length)"""
    from scipy.spatial import cKDTree
    Lb = np.array([c.lat[0], c.lat[4], c.lat[8]])
    assert rc < Lb.min()
    shifts = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], dtype=np.float64)
    pts = (c.pos.T[None, :, :] + (shifts * Lb)[:, None, :]).reshape(-1, 3)
    owner = np.tile(np.arange(c.n), 27)
    tree = cKDTree(pts)
    ordered, selfimg = 0, 0
    for i, nb in enumerate(tree.query_ball_point(c.pos.T, rc)):
        nb = np.asarray(nb)
        ordered += nb.size - 1                               # minus (i, no shift)
        selfimg += int((owner[nb] == i).sum()) - 1
    return (ordered - selfimg) // 2 + selfimg


def c2_section(L, W, args):
    """Config 2, the ionic path: the rock-salt cell (510 ions, charges +-2,
    Buckingham, Ewald eta = 0.35 1/A, rc = 10 A over explicit images, |n| < 8,
    1147 k vectors).  Per-pass device times of one cell and of the 16-walker
    batch (HIP events, time_phases), force evaluations per second, the
    real-space roofline (45 FLOP per unordered in-cutoff pair, images
    included, FP32 pair terms) and the reciprocal one (20 N K FP64 FLOP),
    then 16 batched ART searches from seeded starts with device FP64
    verification (ART saddles/s)."""
    c = W.rocksalt_cell(seed=5)
    pots = [("M", "X") + W.ROCKSALT_BUCK[("M", "X")], ("X", "X") + W.ROCKSALT_BUCK[("X", "X")],
            ("M", "M") + W.ROCKSALT_BUCK[("M", "M")]]
    f = L.Field(L.buck_potentials_text(["M", "X"], [2.0, -2.0], pots, 10.0, noimage=True))
    f.set_ewald(0.35, 8)
    nrep = 200
    b1 = L.Batch(f, [c])
    b1.forces()
    K = f.num_kvectors
    ph1 = b1.time_phases(nrep)  # ms per evaluation
    Pc = ion_pairs_with_images(c, 10.0)
    real_flop, recip_flop = 45.0 * Pc, 20.0 * c.n * K
    out = {"cell": {"ions": c.n, "k_vectors": K, "unordered_pairs_with_images": int(Pc),
                    "ms_lists": round(float(ph1[0]), 5), "ms_real_space": round(float(ph1[1]), 5),
                    "ms_reciprocal": round(float(ph1[2]), 5), "ms_reduce": round(float(ph1[3]), 5),
                    "force_evals_per_s_device": round(1e3 / float(ph1.sum()), 1),
                    "real_space_tflops": round(real_flop / (ph1[1] * 1e-3) / 1e12, 4),
                    "reciprocal_fp64_tflops": round(recip_flop / (ph1[2] * 1e-3) / 1e12, 4)}}
    b1.close()
    nw = 16
    b = L.Batch(f, [c] * nw)
    b.forces()
    ph = b.time_phases(50)
    out["batch16"] = {"ms_real_space": round(float(ph[1]), 5), "ms_reciprocal": round(float(ph[2]), 5),
                      "ms_total": round(float(ph.sum()), 5),
                      "force_evals_per_s_cells_device": round(nw * 1e3 / float(ph.sum()), 1)}
    out["roofline_real_space"] = {"bound": "valu", "kernel": "k_ion_real (erfc real space + Buckingham)",
                                  "achieved": round(nw * real_flop / (ph[1] * 1e-3) / 1e12, 4), "unit": "TFLOP/s",
                                  "peak": FP32_VECTOR_PEAK_TFLOPS,
                                  "frac": round(nw * real_flop / (ph[1] * 1e-3) / 1e12 / FP32_VECTOR_PEAK_TFLOPS, 5),
                                  "flops_per_launch": nw * real_flop, "ms_per_launch": round(float(ph[1]), 5)}
    out["roofline_reciprocal"] = {"bound": "valu", "kernel": "k_ewald_tables + k_ewald_sk + k_ewald_force (FP64)",
                                  "achieved": round(nw * recip_flop / (ph[2] * 1e-3) / 1e12, 4), "unit": "TFLOP/s",
                                  "peak": FP64_VECTOR_PEAK_TFLOPS,
                                  "frac": round(nw * recip_flop / (ph[2] * 1e-3) / 1e12 / FP64_VECTOR_PEAK_TFLOPS, 5),
                                  "flops_per_launch": nw * recip_flop, "ms_per_pass": round(float(ph[2]), 5)}
    if args.c2_no_art:
        b.close()
        f.close()
        return out
    # 16 batched ART searches with device FP64 verification
    b.set_frozen([np.zeros(c.n, np.int32)] * nw)
    b.start_relax(walkers=[0])
    b.run(20000, sync_every=50)
    basin = b.positions()[0]
    cb = W.Cell(basin, c.types, c.lat, charges=c.charges)
    b.upload([cb] * nw)
    b.save_basin()
    e32, e64 = b.search_energies(32), b.search_energies(64)
    site = c.vacancies[0]
    d = W.min_image(basin - site[:, None], c.lat)
    r = np.sqrt((d * d).sum(0))
    frozen = (r >= 6.0).astype(np.int32)
    cations = np.nonzero(c.types == 0)[0]
    movers = cations[np.argsort(r[cations])[:12]]
    x0s, modes = [], []
    for w in range(nw):
        disp, mode = W.local_displacement(cb, movers, 0.1, 100 + w)
        x0s.append(basin + disp)
        modes.append(mode)
    b.set_verify(16, 0.01)
    t0 = time.perf_counter()
    b.start_art(modes, params=L.art_defaults(initTimeStep=0.0005, maxTimeStep=0.002, forceTol=1e-2, parDamp=0.25,
                                             numLanczosVectors=25), x0_per_walker=x0s, frozen_per_walker=[frozen] * nw)
    steps = b.run(200000, sync_every=50)
    t_art = time.perf_counter() - t0
    st = b.status()
    conv = [s for s in st if s["status"] == 0]
    ver = [s for w, s in enumerate(st) if s["status"] == 0 and s["verified"] and s["fmax64"] < 0.05
           and abs((s["energy32"] - e32[w]) - (s["energy64"] - e64[w])) <= 0.01 and s["lambda1"] < 0 < s["lambda2"]]
    out["art16"] = {"searches": nw, "converged": len(conv), "fp64_verified": len(ver), "seconds": round(t_art, 3),
                    "micro_steps": steps, "force_calls": int(sum(s["force_calls"] for s in st)),
                    "verified_saddles_per_s": round(len(ver) / t_art, 2),
                    "barriers_eV": sorted(round(s["energy64"] - e64[w], 4) for w, s in enumerate(st)
                                          if s["status"] == 0 and s["verified"])}
    b.close()
    f.close()
    if not args.no_cpu:
        from oracle import oracle as O
        ew = O.Ewald(c.lat, 0.35, gcell=8)
        of = O.ion_field(["M", "X"], pots, rc=10.0, ewald=ew, min_image=False, lat=c.lat)
        ts = []
        for k in range(13):
            t1 = time.perf_counter()
            O.ion_forces(of, c.pos, c.types, c.charges, c.lat)
            if k >= 3:
                ts.append(time.perf_counter() - t1)
        med = float(np.median(ts))
        out["cpu_baseline"] = {"kind": "port", "cores": 1, "cpu_model": cpu_model(),
                               "sample": "serial FP64 Buckingham + Ewald force evaluations of the same cell "
                                         "(3 untimed, median of 10)",
                               "median_ms": round(1e3 * med, 3), "force_evals_per_s": round(1.0 / med, 2),
                               "gpu_cell_over_cpu": round((1e3 / float(ph1.sum())) / (1.0 / med), 1)}
    return out


def kmc_c5_section(L, W, args):
    """The full AKMC loop on one GPU: the 31,980-atom Cu/Ni Sutton-Chen alloy
    (20 x 20 x 20 FCC, 50/50 seeded occupancy, 20 vacancies), T = 600 K,
    prefactor 1e13 1/s, each KMC step's catalogue rebuilt by --c5-walkers
    batched searches on local search cells (the 3600 atoms nearest each
    vacancy), learned proposals for half of the walkers once the network has
    been trained (after the first step) with a random-displacement fallback,
    every saddle FP64-verified and followed by its device basin-2 stage;
    --c5-steps KMC steps of the 1000 the workload names."""
    from akmc_amd import kmc_cluster as KC
    from akmc_amd.kmc import KMCParams
    cell = W.alloy_cell(seed=7)
    p = KMCParams(kmcTemperature=600.0, kmcPreFactor=1.0e13)
    t0 = time.perf_counter()
    run = KC.ClusterKMC(W.alloy_potentials_text(), cell, walkers=args.c5_walkers, cluster_atoms=3600, params=p,
                        seed=args.seed, learned_fraction=0.5, skin=args.skin)
    t_init = time.perf_counter() - t0
    st = run.run(args.c5_steps)
    steps = st.kmc_steps
    per_step = [r.seconds for r in st.steps]
    out = {"atoms": cell.n, "vacancies": len(cell.vacancies), "species": "Cu/Ni 50/50", "temperature_K": 600.0,
           "prefactor_per_s": 1.0e13, "walkers": args.c5_walkers, "search_cell_atoms": run.K,
           "search_cell_r_in_A": round(float(min(run.r_in)), 3), "kmc_steps": steps,
           "kmc_steps_of_workload": 1000, "seconds": round(st.seconds, 2), "init_seconds": round(t_init, 2),
           "kmc_steps_per_s": round(steps / st.seconds, 4),
           "seconds_per_kmc_step": round(st.seconds / max(steps, 1), 2),
           "projected_seconds_1000_steps": round(1000 * st.seconds / max(steps, 1), 0),
           "valid_events_per_s": round(st.valid_events / st.seconds, 2),
           "searches": st.searches, "learned_searches": st.learned_searches,
           "fallback_searches": st.fallback_searches, "converged": st.converged, "verified": st.verified,
           "new_basin": st.new_basin, "valid_events": st.valid_events, "learned_events": st.learned_events,
           "learned_share_of_valid_events": round(st.learned_events / max(st.valid_events, 1), 3),
           "out_of_window": st.out_of_window, "outside_search_cell": st.outside_cell, "cancelled": st.cancelled,
           "force_calls_per_valid_event": round(st.force_calls / max(st.valid_events, 1), 1),
           "micro_steps": st.micro_steps, "training_rows": st.training_rows,
           "time_split_s": {"basin1_relax": round(st.t_relax, 2), "search_cells": round(st.t_cells, 2),
                            "searches": round(st.t_search, 2), "training": round(st.t_train, 2)},
           "per_step": [{"events": r.events, "learned_events": r.learned_events, "Ea_eV": round(r.activation_energy, 4),
                         "dt_s": r.dt, "seconds": r.seconds} for r in st.steps],
           "simulated_time_s": run.time, "basin1_energy_eV": round(run.e_full, 6)}
    run.batch.close()
    run.full.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--window", type=int, default=500, help="micro-steps per step")
    ap.add_argument("--sync-every", type=int, default=25,
                    help="micro-steps per device chunk between host checks (refills)")
    ap.add_argument("--walkers", type=int, default=512)
    ap.add_argument("--roofline-window", type=int, default=100,
                    help="micro-steps of the eagerly launched window that times the dominant kernel")
    ap.add_argument("--plain", type=int, default=32, help="plain GPU dimer searches")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skin", type=float, default=None, help="Verlet skin of the device lists (A)")
    ap.add_argument("--no-pipeline", action="store_true",
                    help="windows wait for the refills before the next chunk (no chunk launched ahead)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal section")
    ap.add_argument("--train-epochs", type=int, default=20)
    ap.add_argument("--train-saddles", type=int, default=50000,
                    help="FP64-verified saddles of the campaign the MLP is trained on")
    ap.add_argument("--no-kmc", action="store_true", help="skip the KMC-step section")
    ap.add_argument("--no-eam", action="store_true", help="skip the tabulated-EAM section")
    ap.add_argument("--no-c5", action="store_true", help="skip the full-loop (31,980-atom alloy) section")
    ap.add_argument("--no-c2", action="store_true", help="skip the ionic (rock-salt Ewald) section")
    ap.add_argument("--c2-only", action="store_true", help="only the ionic section (for kernel traces)")
    ap.add_argument("--c2-no-art", action="store_true", help="ionic section without the ART searches")
    ap.add_argument("--c5-steps", type=int, default=3, help="KMC steps of the full-loop section")
    ap.add_argument("--c5-walkers", type=int, default=1024)
    ap.add_argument("--kmc-systems", type=int, default=64)
    ap.add_argument("--kmc-steps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--searches-per-basin", type=int, default=5,
                    help="searches on a basin before the walker moves to its next pooled basin (0: never)")
    ap.add_argument("--basin-pool", type=int, default=8, help="pre-relaxed basins per walker")
    ap.add_argument("--eager", action="store_true",
                    help="launch the micro-step kernels one by one instead of replaying hipGraphs "
                         "(for kernel tracers that cannot follow graph launches)")
    ap.add_argument("--forces-only", type=int, default=0, metavar="N",
                    help="only N force evaluations of all walkers (for hardware-counter runs)")
    ap.add_argument("--zone-passes", type=int, default=0, metavar="N",
                    help="only N evaluations of running searches as the micro-steps run them (counter runs)")
    ap.add_argument("--copy-mib", type=int, default=0, metavar="MIB",
                    help="only a plain device copy of MIB MiB, 5 times (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the walkers of this path "
              "would shard across GPUs with no exchange during a search", file=sys.stderr)
        sys.exit(2)

    if args.copy_mib:
        copy_calibration(args.copy_mib)
        return
    from akmc_amd import lib as L
    from akmc_amd import search as S
    from akmc_amd import workloads as W
    L.load()
    if args.forces_only:
        forces_only(L, W, args)
        return
    if args.zone_passes:
        zone_passes(L, S, W, args)
        return
    if args.c2_only:
        print(json.dumps({"c2": c2_section(L, W, args)}), flush=True)
        return
    cu = W.SC_CU
    rc = 2 * cu["a"]
    text = L.sc_potentials_text([("Cu", "Cu", cu["eps"], cu["a"], cu["n"], cu["m"], cu["c"])], rc)
    cells = W.batched_sc_cells(n_walkers=args.walkers, seed=args.seed)
    natoms = np.array([c.n for c in cells])
    dimer = L.dimer_defaults(maxParIter=400)
    camp = S.SaddleCampaign(text, cells, params=dimer, seed=args.seed, skin=args.skin)
    if args.eager:
        camp.batch.set_graphs(False)
    # kernel variant chosen on the seeded cells, before the relaxation that uses it
    camp.batch.forces()
    variant, variant_ms = camp.batch.autotune(10)
    print(f"[bench] autotune (lanes, unroll, staged): {variant} (ms per evaluation "
          f"{ {k: round(v, 3) for k, v in variant_ms.items()} })", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ok = camp.relax_basins()
    t_relax = time.perf_counter() - t0
    print(f"[bench] {len(cells)} walkers, {natoms.sum()} atoms; relaxed {ok.sum()} basins in {t_relax:.1f} s",
          file=sys.stderr, flush=True)
    if args.searches_per_basin > 0 and args.basin_pool > 1:
        t0 = time.perf_counter()
        camp.relax_pool(args.basin_pool, seed=args.seed + 17)
        camp.searches_per_basin = args.searches_per_basin
        print(f"[bench] basin pool: {args.basin_pool} basins per walker relaxed in {time.perf_counter() - t0:.1f} s; "
              f"a walker moves on after {args.searches_per_basin} searches", file=sys.stderr, flush=True)

    # ---- roofline of the force kernels at the basins (lists settled)
    camp.batch.upload([W.Cell(p, c.types, c.lat) for p, c in zip(camp.basins, cells)])
    camp.batch.forces()
    ph = camp.batch.time_phases(20)
    P = ordered_pairs(cells, camp.basins, rc)
    Ptot = int(P.sum())
    A = int(natoms.sum())
    force_ms = ph[2]
    achieved_tf = 40.0 * Ptot / (force_ms * 1e-3) / 1e12
    dens_tf = 20.0 * Ptot / (ph[1] * 1e-3) / 1e12
    eval_ms = float(ph.sum())
    hbm_gbs = 56.0 * A / ((ph[1] + ph[2]) * 1e-3) / 1e9
    print(f"[bench] phases ms: lists {ph[0]:.3f} density {ph[1]:.3f} force {ph[2]:.3f} reduce {ph[3]:.3f}; "
          f"P={Ptot}", file=sys.stderr, flush=True)

    # ---- steady-state windows of batched searches with refill
    camp.begin()
    tot = S.CampaignStats()
    t_timed = 0.0
    search_variant = None
    for r in range(args.warmup + args.steps):
        if r == 1 or (r == 0 and args.warmup == 0):
            # the search micro-steps' own variants, tuned per pass on the
            # searches under way (zoned passes): untimed, between windows
            t0 = time.perf_counter()
            camp.drain()
            sv, sv_ms = camp.batch.autotune_search(5)
            search_variant = {"pass1": sv[:3], "pass2": sv[3:6], "fp64": sv[6:9],
                              "seconds": round(time.perf_counter() - t0, 2),
                              "ms_per_5_evals_pass1_pass2_fp64":
                                  {f"{k[0]}x{k[1]}{' lds' if k[2] else ''}":
                                   [round(x, 4) if x is not None else None for x in v] for k, v in sv_ms.items()}}
            print(f"[bench] search variants (lanes, unroll, staged): pass 1 {sv[:3]} pass 2 {sv[3:6]} "
                  f"FP64 {sv[6:9]}; {search_variant['ms_per_5_evals_pass1_pass2_fp64']}", file=sys.stderr,
                  flush=True)
        rb0 = camp.batch.rebuilds
        st = camp.advance(args.window, sync_every=args.sync_every, pipelined=not args.no_pipeline)
        print(f"[bench] window {r}: {camp.batch.rebuilds - rb0} walker list rebuilds; "
              f"{st.searches} ended, {st.converged} converged, {st.verified} verified, "
              f"{st.distinct} distinct, {st.micro_steps} micro-steps in {st.search_time:.2f}s, "
              f"max barrier err {st.max_barrier_err:.2e}; failed tests {st.failures()}; device {st.device_time:.2f}s "
              f"harvest {st.harvest_time:.2f}s refill {st.refill_time:.2f}s; per chunk part: launch {st.t_launch:.3f} "
              f"overlapped host {st.t_overlap:.3f} wait {st.t_wait:.3f} collect {st.t_collect:.3f} "
              f"start {st.t_start:.3f} s", file=sys.stderr, flush=True)
        if r >= args.warmup:
            t_timed += st.search_time
            tot.add(st)
    # ---- the dominant kernel timed with HIP events on the batch's stream over a
    # window of the same searches that follows the timed ones (events inside
    # replayed graphs cannot be timed, so this window launches eagerly), with
    # the FP32 evaluations each walker consumed there
    camp.drain()
    counts0 = camp.batch.eval_counts()
    camp.batch.kernel_timer(True)
    rst = camp.advance(args.roofline_window, sync_every=25)
    k_ms, k_launches = camp.batch.kernel_time()
    evals = camp.batch.eval_counts() - counts0
    camp.batch.kernel_timer(False)
    timed_ms = k_ms / max(k_launches, 1)
    print(f"[bench] k_sc_force in the roofline window: {k_launches} launches, {timed_ms:.4f} ms per launch "
          f"({int(evals.sum())} walker evaluations)", file=sys.stderr, flush=True)
    # ---- the pass kernels as the search micro-steps run them (active zones),
    # timed with HIP events on the batch's stream at the searches' current
    # probes, after one warm evaluation that does any pending full one; the
    # algorithmic work from the device's zones and the free-atom masks
    camp.batch.time_search_phases(1)
    zw = zone_work(camp, cells, rc)
    zph = camp.batch.time_search_phases(20)
    z_tf1 = ZONE_PASS1_FLOP * zw["pairs_pass1"] / (zph[1] * 1e-3) / 1e12
    z_tf2 = 40.0 * zw["pairs_pass2"] / (zph[2] * 1e-3) / 1e12
    print(f"[bench] zoned search evaluation: {zw}; ms lists {zph[0]:.4f} pass 1 {zph[1]:.4f} ({z_tf1:.2f} TFLOP/s) "
          f"pass 2 {zph[2]:.4f} ({z_tf2:.2f} TFLOP/s) reduce {zph[3]:.4f}", file=sys.stderr, flush=True)
    value = tot.distinct / t_timed
    verified_rate = tot.verified / t_timed
    moves_timed = camp.basin_moves
    batched_force_evals = tot.force_calls / t_timed

    # ---- plain GPU dimer on the same cells
    n_ok, t_plain, plain_calls = plain_dimer_rate(L, W, text, cells, camp.basins, camp.proposal, dimer,
                                                  args.plain, camp.basin32, camp.basin64, eager=args.eager)
    plain_rate = n_ok / t_plain if t_plain > 0 else 0.0  # 32 searches on 32 cells: every verified one is distinct
    print(f"[bench] plain GPU dimer: {n_ok}/{args.plain} converged in {t_plain:.2f}s "
          f"({plain_calls} force calls)", file=sys.stderr, flush=True)

    learned = None
    if not args.no_learned:
        learned = learned_ab(camp, tot.records, args.window, args.train_epochs, args.train_saddles)
        print(f"[bench] learned section: { {k: v for k, v in learned.items() if k not in ('random', 'learned')} }",
              file=sys.stderr, flush=True)

    eam = None
    if not args.no_eam:
        eam = eam_section(L, W, cells, camp.basins, P, args)
        print(f"[bench] tabulated EAM: {eam}", file=sys.stderr, flush=True)

    kmc = None
    if not args.no_kmc:
        kmc = kmc_section(L, W, text, args)
        print(f"[bench] KMC steps: {kmc}", file=sys.stderr, flush=True)

    c2 = None
    if not args.no_c2:
        c2 = c2_section(L, W, args)
        print(f"[bench] ionic section: {c2}", file=sys.stderr, flush=True)

    c5 = None
    if not args.no_c5:
        c5 = kmc_c5_section(L, W, args)
        print(f"[bench] full loop (alloy): { {k: v for k, v in c5.items() if k != 'per_step'} }", file=sys.stderr,
              flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cells, camp.basins, camp.proposal)
        print(f"[bench] CPU baseline: {cpu}", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "saddles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_timed / args.steps, 3),
        "micro_steps_per_step": args.window,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair terms, fp64 cutoff/accumulation",
        "data": "synthetic (seeded FCC Sutton-Chen Cu cells with vacancies)",
        "config": {"workload": "batched dimer searches: 512 Sutton-Chen Cu walkers of 500-4000 atoms, "
                               "1-3 vacancies, FP64-verified", "walkers": len(cells), "atoms": A,
                   "parallelism": "walkers batched on one GPU"},
        "verified_saddles_per_s": round(verified_rate, 3),
        "searches": tot.searches, "converged": tot.converged, "verified": tot.verified,
        "distinct": tot.distinct, "max_barrier_error_eV": tot.max_barrier_err,
        "verification_failures": tot.failures(),
        "plain_gpu_dimer_saddles_per_s": round(plain_rate, 3),
        "plain_gpu_dimer": {"searches": args.plain, "verified": n_ok, "seconds": round(t_plain, 3),
                            "force_calls_per_search": round(plain_calls / max(1, args.plain), 1),
                            "ms_per_force_call": round(1e3 * t_plain / max(1, plain_calls), 4)},
        "batched_force_calls_per_verified": round(tot.force_calls / max(1, tot.verified), 1),
        "windows": {"pipelined": not args.no_pipeline, "micro_steps_per_chunk": args.sync_every},
        "basin_turnover": {"searches_per_basin": args.searches_per_basin, "pool_basins_per_walker": args.basin_pool,
                           "basin_moves": int(moves_timed)},
        "batched_over_plain": round(value / plain_rate, 2) if plain_rate > 0 else None,
        "batched_over_plain_basis": "FP64-verified distinct saddles/s, batched (value) over the plain GPU dimer",
        "batched_over_plain_verified": round(verified_rate / plain_rate, 2) if plain_rate > 0 else None,
        "force_evals_per_s_batched_cells": round(batched_force_evals, 1),
        "force_eval_ms_all_walkers": round(eval_ms, 4),
        "force_evals_per_s_cells_device": round(len(cells) / (eval_ms * 1e-3), 1),
        "search_variant": search_variant,
        "autotune": {"lanes_per_atom": variant[0], "gathers_in_flight": variant[1], "lds_staged": bool(variant[2]),
                     "ms_per_eval": {f"{k[0]}x{k[1]}{' lds' if k[2] else ''}": round(v, 4)
                                     for k, v in variant_ms.items()}},
        "roofline": {"bound": "valu", "bound_note": "compute-bound on the FP32 vector ALU (the pair physics has "
                     "no matrix form and uses no MFMA); peak is MI355X's published FP32 vector figure (packed "
                     "FP32). This build compiles without packed FP32 instructions, so its reachable FP32 "
                     "vector ceiling is half that: frac_vs_unpacked_ceiling",
                     "kernel": "k_sc_density, zoned (Sutton-Chen pass 1 of the search micro-steps: density and "
                               "pair energy of each walker's active zone, autotuned variant)",
                     "achieved": round(z_tf1, 3), "peak": FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(z_tf1 / FP32_VECTOR_PEAK_TFLOPS, 4),
                     "unpacked_ceiling": FP32_VECTOR_PEAK_TFLOPS / 2,
                     "frac_vs_unpacked_ceiling": round(z_tf1 / (FP32_VECTOR_PEAK_TFLOPS / 2), 4),
                     "traffic": round(ZONE_TRAFFIC_BYTES_PER_PAIR * zw["pairs_pass1"]),
                     "traffic_note": "HBM bytes of one zoned pass-1 launch: 9.72 B per ordered pair (FETCH_SIZE x "
                                     "2.0 + WRITE_SIZE x 1.0, scales from a 1 GiB copy) from the counter runs of "
                                     "bench.py --zone-passes 10, times this launch's pairs",
                     "measured_over": "20 evaluations of the search walkers at their probes after the timed "
                                      "windows, as the micro-steps launch them (HIP events on the batch's stream)",
                     "ms_per_launch": round(float(zph[1]), 4),
                     "flops_per_launch": round(ZONE_PASS1_FLOP * zw["pairs_pass1"]),
                     "flop_per_pair": ZONE_PASS1_FLOP,
                     "zones": zw,
                     "pass2_zoned": {"kernel": "k_sc_force (free atoms)", "ms_per_launch": round(float(zph[2]), 4),
                                     "flops_per_launch": 40 * zw["pairs_pass2"], "tflops": round(z_tf2, 3)},
                     "window_k_sc_force_ms_per_launch": round(timed_ms, 4),
                     "full_evaluation": {"kernel": "k_sc_force over every atom of every walker at the basins",
                                         "ms_per_launch": round(force_ms, 4), "flops_per_launch": 40 * Ptot,
                                         "tflops": round(achieved_tf, 3),
                                         "frac": round(achieved_tf / FP32_VECTOR_PEAK_TFLOPS, 4),
                                         "density_pass_tflops": round(dens_tf, 3),
                                         "traffic": round(TRAFFIC_BYTES_PER_PAIR * Ptot) if TRAFFIC_BYTES_PER_PAIR
                                         else None,
                                         "traffic_note": "FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (scales from a "
                                                         "1 GiB copy), from the hardware-counter run",
                                         "hbm_algorithmic_GBs": round(hbm_gbs, 1)},
                     "hbm_peak_GBs": HBM_PEAK_GBS},
    }
    if learned:
        out["learned_proposal"] = learned
    if kmc:
        out["kmc"] = kmc
    if eam:
        out["tabulated_eam"] = eam
    if c5:
        out["kmc_c5"] = c5
    if c2:
        out["ionic_c2"] = c2
    if cpu:
        out["cpu_baseline"] = {"value": round(cpu["saddles_per_s"], 5), "unit": "saddles/s",
                               "cores": cpu["cores"], "kind": "port", "cpu_model": cpu["model"],
                               "sample": f"{cpu['searches']} dimer searches on cells of {cpu['sizes']} atoms "
                                         f"({cpu['seconds']:.1f} s, O(N^2) FP64 Sutton-Chen, OpenMP)",
                               "force_evals_per_s": round(cpu["force_evals_per_s"], 2),
                               "force_eval_timing": cpu["force_timing"]}
        out["gpu_over_cpu_saddles"] = round(verified_rate / cpu["saddles_per_s"], 1) if cpu["saddles_per_s"] else None
        out["plain_gpu_over_cpu_saddles"] = round(plain_rate / cpu["saddles_per_s"], 1) if cpu["saddles_per_s"] else None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
