"""Benchmark of the pair-distance / pair-histogram hot path on one MI355X.

Headline workload (N=1): the SiO2-like box of 200,000 atoms (66,667 Si,
133,333 O, uniform in a 145 A cube with a 1.5 A minimum distance), the
largest configuration of the benchmark set.  One O(N^2) full histogram on
r in [0, 20) at 0.01 A (2,000 bins; float32 reference binning and the
binary64-recompute mode), then single-atom translation moves of 0.1 A with a
PairDistributionConstraint (the same 2,000 bins) and a
StructureFactorConstraint (own r grid [0, 20) at 0.01 A, Q = 0.3..30 1/A,
1,000 points), windows of up to 128 speculative proposals.  The targets are
the float64 G(r) / S(Q) of an independent box of the same composition.

One step = ``--moves`` generated moves (default 5,000) through the
device-resident engine; the warm-up steps are the run's fixed
pre-equilibration, and the timed steps (default 20) are the 100,000 moves of
the workload.  ``value`` = generated moves per second over the timed steps,
with every input resident in HBM.  The trajectory of the first timed step is
replayed by the CPU oracle from the state at the start of the timed region
(same random streams) for the CPU baseline, and its moves are checked
against the device's (``parity_checked_moves``).

Also reported: the plain one-proposal-per-window form on the same box; the
13,500-atom B2 alloy (G(r) + S(Q)), the 256 batched replicas and the 6,500-
atom molecular triclinic box with their own CPU baselines; swap moves and
removals on the B2 box and the molecular box with a distance constraint,
each beside the reference's per-move loop over the drop-in calls; the S(Q)
transform entry; the roofline of the dominant kernel.
"""
import argparse
import ctypes
import json
import math
import os
import random
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X dense FP32 (vector and matrix)
PEAK_HBM_GBS = 8000.0
OPS_PER_PAIR = 35          # PBC pair: diff 3, round 6x3, sub 3, matvec 15, norm2 5, sqrt 1, bin 2 (minus 9 -> 35)


def target_curves(start, target, dr, H, q):
    """float64 G(r) on r = dr/2 + dr*k (k < H) and S(Q) on q of the target
    box, from its pair counts (GPU full histogram), with the start box's
    composition; returns (elem of start, pdc experimentalData, sfc
    experimentalData) as float32."""
    from fullrmc_amd import systems
    from fullrmc_amd.Constraints.PairDistributionConstraints import ATOMIC_NUMBER, normalized_weighting
    from fullrmc_amd.Core.pairs_histograms import full_pairs_histograms_coords

    els = sorted(set(start["elements"]))
    lut = {e: i for i, e in enumerate(els)}
    elem = np.array([lut[e] for e in start["elements"]], dtype=np.int32)
    telem = np.array([lut[e] for e in target["elements"]], dtype=np.int32)
    nel = len(els)
    counts = {e: start["elements"].count(e) for e in els}
    n = len(elem)
    V = float(systems.volume(start["basis"]))
    rho = n / V
    w = {k: float(v) for k, v in normalized_weighting(counts, {e: ATOMIC_NUMBER[e] for e in els}).items()}
    tbox = systems.box_coordinates(target["recip"], target["real"])
    hi, he = full_pairs_histograms_coords(tbox, target["basis"], True, target["molecules"], telem, nel,
                                          np.float32(0.0), np.float32(dr * H), np.float32(dr), H)
    r = dr / 2 + dr * np.arange(H)
    edges = dr * np.arange(H + 1)
    shell = 4.0 / 3.0 * np.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    g = np.zeros(H)
    for a in range(nel):
        for b in range(a, nel):
            ea, eb = els[a], els[b]
            if a == b:
                nij = hi[a, a].astype(np.float64) + he[a, a]
                D = counts[ea] * (counts[ea] - 1) / 2.0 / V
            else:
                nij = (hi[a, b].astype(np.float64) + hi[b, a] + he[a, b] + he[b, a])
                D = counts[ea] * counts[eb] / V
            g += w[ea + "-" + eb] * nij / D / shell
    G = 4 * np.pi * r * rho * (g - 1)
    S = 1 + np.sum(G[:, None] * dr * np.sin(np.outer(r, q)) / q, axis=0)
    pdc = np.stack([r, G], 1).astype(np.float32)
    sfc = np.stack([q, S], 1).astype(np.float32)
    return elem, pdc, sfc


def build_workload(seed=2026):
    """B2 alloy, 13,500 atoms (15x15x30 cells), PDC r in [0,20) @0.02, SFC on
    its own r grid with Q = linspace(0.5, 25, 500); the target is an
    independently jittered copy."""
    from fullrmc_amd import systems
    start = systems.b2_alloy(cells=(15, 15, 30), a=3.01, sigma=0.08, seed=seed)
    target = systems.b2_alloy(cells=(15, 15, 30), a=3.01, sigma=0.08, seed=seed + 1)
    elem, pdc, sfc = target_curves(start, target, 0.02, 1000, np.linspace(0.5, 25.0, 500))
    return start, elem, pdc, sfc


def build_sio2_workload(seed=2030):
    """SiO2-like box, 200,000 atoms (66,667 Si) in a 145 A cube with a 1.5 A
    minimum distance; PDC r in [0,20) @0.01 (2000 bins), SFC on its own r grid
    with Q = linspace(0.3, 30, 1000); the target is an independent box."""
    from fullrmc_amd import systems
    start = systems.sio2_box(200000, 145.0, 1.5, seed=seed)
    target = systems.sio2_box(200000, 145.0, 1.5, seed=seed + 1)
    target["elements"] = start["elements"]  # same composition and species order
    elem, pdc, sfc = target_curves(start, target, 0.01, 2000, np.linspace(0.3, 30.0, 1000))
    return start, elem, pdc, sfc


def make_engine(start, pdc, sfc, seed, dr=0.02):
    from fullrmc_amd.Constraints import PairDistributionConstraint, StructureFactorConstraint
    from fullrmc_amd.Engine import Engine
    from fullrmc_amd.Generators import TranslationGenerator
    random.seed(seed)
    np.random.seed(seed)
    E = Engine()
    E.set_system(start["real"], start["elements"], start["molecules"], start["basis"], start["recip"])
    E.set_move_generator(TranslationGenerator(0.1))
    E.add_constraints([PairDistributionConstraint(pdc),
                       StructureFactorConstraint(sfc, rmin=0.0, rmax=20.0, dr=dr)])
    return E


def gpu_full_histogram(start, elem, dr, H, reps=3, float64=False, autotune=True):
    """One O(N^2) recompute on the PDC grid (r in [0, dr*H)): median kernel
    time by HIP events; float64=True times the binary64-recompute mode and
    also returns its number of double re-checks.  autotune=False times the
    fixed-rule schedule instead of the tuned one; the schedule used is
    returned last."""
    from fullrmc_amd import _device as D
    from fullrmc_amd import _lib, systems, tuning
    L = _lib.lib()
    box = systems.box_coordinates(start["recip"], start["real"])
    n = len(box)
    nel = int(elem.max()) + 1
    d_c, d_b = D.upload(box), D.upload(start["basis"])
    d_m, d_e = D.upload(start["molecules"]), D.upload(elem)
    counts = D.empty((2 * nel * nel * H + 4,), np.int32)
    rechecked = ctypes.c_int64(0)
    tuning.set_autotune(autotune)
    L.frmc_set_timing(1)
    ms = []
    for _ in range(reps + 1):
        if float64:
            _lib.check(L.frmc_full_pairs_counts_f64(d_c.data_ptr(), n, d_b.data_ptr(), 1, d_m.data_ptr(),
                                                    d_e.data_ptr(), nel, 0.0, dr * H, dr, H, counts.data_ptr(),
                                                    ctypes.byref(rechecked), D.stream()))
        else:
            _lib.check(L.frmc_full_pairs_counts(d_c.data_ptr(), n, d_b.data_ptr(), 1, d_m.data_ptr(),
                                                d_e.data_ptr(), nel, 0.0, dr * H, dr, H, counts.data_ptr(),
                                                D.stream()))
        ms.append(L.frmc_last_kernel_ms())
    evaluated = ctypes.c_int64(-1)
    _lib.check(L.frmc_last_full_evaluated(ctypes.byref(evaluated)))
    L.frmc_set_timing(0)
    sched = tuning.last_full_schedule()
    sched["evaluated_pairs"] = int(evaluated.value)
    tuning.set_autotune(True)
    ms = sorted(ms[1:])[len(ms[1:]) // 2]
    pairs = n * (n - 1) / 2
    return pairs / (ms * 1e-3), ms, int(rechecked.value), sched


def cpu_engine(start, pdc, sfc, dr, seed=7, init_from=None):
    """The oracle engine on the workload; init_from (a snapshot of a device
    engine, see snapshot()) supplies the coordinates and slot-exact
    histograms instead of an O(N^2) oracle recompute, and the oracle's
    total must then equal the snapshot's."""
    from oracle import chain
    from oracle import engine as oeng
    from fullrmc_amd.Constraints.PairDistributionConstraints import ATOMIC_NUMBER
    random.seed(seed)
    np.random.seed(seed)
    real = start["real"] if init_from is None else init_from["real"]
    osys = oeng.System(real, start["elements"], start["molecules"], start["basis"], start["recip"])
    w = chain.normalized_weighting(osys.counts, {e: ATOMIC_NUMBER[e] for e in osys.elements})
    oe = oeng.Engine(osys, [[i] for i in range(len(start["real"]))], oeng.Translation(0.1),
                     [oeng.PDC(osys, pdc, w), oeng.SFC(osys, sfc, w, 0.0, 20.0, dr)])
    if init_from is None:
        oe.initialize()
        return oe
    assert osys.box.tobytes() == init_from["box"].tobytes()
    for c, (hi, he) in zip(oe.constraints, init_from["hist"]):
        c.intra, c.inter = np.array(hi, np.float32), np.array(he, np.float32)
        c.sd = c.chi2(c.intra, c.inter)
    oe.tse = chain.total_error([c.sd for c in oe.constraints])
    assert np.float32(oe.tse).tobytes() == np.float32(init_from["tse"]).tobytes()
    return oe


def snapshot(E):
    """The device engine's state (coordinates, histograms, total) and both
    global random streams, for an oracle replay."""
    E.sync_from_device()
    return {"real": np.array(E.realCoordinates, np.float32), "box": np.array(E.boxCoordinates, np.float32),
            "hist": [(np.array(c.data["intra"]), np.array(c.data["inter"])) for c in E.constraints],
            "tse": np.float32(E.totalStandardError), "py": random.getstate(), "np": np.random.get_state()}


def cpu_moves_per_s(oe, seconds):
    n = 0
    t0 = time.perf_counter()
    while True:
        oe.run(5)
        n += 5
        el = time.perf_counter() - t0
        if el > seconds:
            return n / el, n, el


def cpu_full_histogram_sample(start, elem, dr, H, rows, threads):
    """The oracle's all-pairs loop (C, OpenMP, private bins) over a sample of
    `rows` evenly spaced row atoms (each against all later atoms): pairs/s."""
    from oracle import refcpu
    from fullrmc_amd import systems
    box = systems.box_coordinates(start["recip"], start["real"])
    n = len(box)
    idx = np.linspace(0, n - 2, rows).astype(np.int32)
    pairs = float(np.sum(n - 1 - idx.astype(np.int64)))
    t0 = time.perf_counter()
    refcpu.multiple_pairs_histograms_coords(idx, box, start["basis"], True, start["molecules"], elem,
                                            int(elem.max()) + 1, np.float32(0.0), np.float32(dr * H),
                                            np.float32(dr), H, allAtoms=False, ncores=threads)
    return pairs / (time.perf_counter() - t0), pairs


# fp32 VALU without FMA and without packed FP32 (the reference's float32
# operation order forbids contraction): one operation per lane per cycle,
# 256 CUs x 64 lanes x 2.4 GHz
ATTAINABLE_FP32_TOPS = 39.3
# a wave's chain of dependent v_add_f32 issues one add per 6.94 cycles
# (measured on MI355X: 2,000 dependent register adds per wave, clock64 around
# the loop), at the 2.4 GHz engine clock
DEPENDENT_ADD_CYCLES = 6.94
CLOCK_GHZ = 2.4


def eval_flops_per_proposal(H_pdc=2000, H_sfc=2000, nQ=1000, T=3):
    """k_eval_a per proposal: G(r) (T multiply-divide-add + divide, subtract,
    multiply per bin), S(Q) (one product and one add per (r, Q)), squared
    deviations (3 per point) and the pairwise sums (1 per point)."""
    gr = (3 * T + 3) * (H_pdc + H_sfc)
    return gr + 2.0 * H_sfc * nQ + 4.0 * (H_pdc + nQ)


def kernel_roofline(name, avg_s, n, props_per_launch, shape=None, traffic=None, traffic_source=None):
    """Algorithmic work of one launch of the named engine kernel against the
    FP32 vector peak: the pair path issues no matrix instructions.  shape:
    (H_pdc, H_sfc, nQ, T) of the evaluation.  traffic: measured HBM bytes per
    launch (rocprofv3 FETCH_SIZE + WRITE_SIZE of the same command), if any."""
    shape = shape or (2000, 2000, 1000, 3)
    if name == "k_delta":
        flops = OPS_PER_PAIR * 2.0 * n * props_per_launch      # old + new position vs all atoms
        limit = "latency of the per-pair atomics into the proposal's delta block; VALU distance arithmetic"
    elif name == "k_eval_a":
        flops = eval_flops_per_proposal(*shape) * props_per_launch
        limit = ("latency: the S(Q) sum of each Q is H dependent float32 adds in the reference's "
                 "sequential order, after the G(r) build")
    else:
        return None
    if not avg_s > 0:
        return None
    tf = flops / avg_s / 1e12
    extra = {}
    if name == "k_eval_a":
        # the sequential floor: each Q point's S(Q) sum is H dependent adds
        floor_us = shape[1] * DEPENDENT_ADD_CYCLES / CLOCK_GHZ * 1e-3
        extra = {"sequential_floor_us": round(floor_us, 2),
                 "frac_of_sequential_floor": round(floor_us / (avg_s * 1e6), 4)}
    return {"bound": "valu", "limited_by": limit,
            "compute": "fp32 VALU, no matrix instructions (bit-exact float32 forbids FMA and packed FP32)",
            "achieved": round(tf, 4), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tf / PEAK_FP32_TFLOPS, 6),
            "attainable_peak": ATTAINABLE_FP32_TOPS,
            "frac_of_attainable": round(tf / ATTAINABLE_FP32_TOPS, 6),
            "traffic": traffic, "traffic_source": traffic_source,
            "kernel": name, "avg_launch_us": round(avg_s * 1e6, 2), "flops_per_launch": flops,
            "proposals_per_launch": round(props_per_launch, 2), **extra}


def committed_traffic(profile, kernel, proposals):
    """HBM bytes per launch of kernel at `proposals` proposals per launch,
    from a committed PMC summary (JSON beside the .md: measured bytes per
    proposal, scaled here), with its source, or (None, None)."""
    try:
        with open(os.path.join(ROOT, "profiles", profile)) as f:
            rec = json.load(f)["kernels"][kernel]
        per = float(rec["traffic_bytes_per_proposal"])
        return int(round(per * proposals)), "profiles/%s (%s; %.0f B per proposal x %d)" % (
            profile, rec["how"], per, proposals)
    except (OSError, KeyError, ValueError, TypeError):
        return None, None


def _replica_boxes(args):
    from fullrmc_amd import systems
    r, n, seed = args
    L = float((n / 0.05) ** (1.0 / 3.0))
    start = systems.random_box(int(n), L, 2.0, seed + 1 + 2 * r)
    target = systems.jitter(start, 0.1, seed + 2 + 2 * r)
    return start, target


def build_replica_boxes(nrep, seed=2040):
    """Independent single-element boxes: N log-uniform in [1000, 20000],
    cubic edge for a number density of 0.05 / A^3, 2.0 A minimum distance,
    and a sigma = 0.1 A jittered copy of each as its target structure.  Built
    by a process pool; call before the GPU is initialised."""
    import multiprocessing
    from fullrmc_amd import systems
    sizes = systems.replica_sizes(nrep, 1000, 20000, seed)
    jobs = [(r, int(n), seed) for r, n in enumerate(sizes)]
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    with multiprocessing.get_context("fork").Pool(workers) as pool:
        return pool.map(_replica_boxes, jobs)


def build_replicas(boxes):
    """Each replica's float64 target G(r) on r in [0, 15) @ 0.05 (GPU pair
    counts of its target structure)."""
    out = []
    for start, target in boxes:
        elem, pdc, _ = target_curves(start, target, 0.05, 300, np.array([1.0]))
        out.append((start, pdc))
    return out


def replicas_section(args, boxes):
    """Independent replicas advanced together (run_replicas), beside the
    oracle engine on a sample of the same replicas."""
    import torch
    from fullrmc_amd.Constraints import PairDistributionConstraint
    from fullrmc_amd.Engine import Engine, run_replicas
    from fullrmc_amd.Generators import TranslationGenerator
    t0 = time.perf_counter()
    reps = build_replicas(boxes)
    engines = []
    for start, pdc in reps:
        E = Engine()
        E.set_system(start["real"], start["elements"], start["molecules"], start["basis"], start["recip"])
        E.set_move_generator(TranslationGenerator(0.1))
        E.add_constraints([PairDistributionConstraint(pdc)])
        engines.append(E)
    seeds = [3000 + r for r in range(len(engines))]
    run_replicas(engines, 20, k=32, seeds=seeds)  # engine creation and warm-up
    torch.cuda.synchronize()
    print("bench: %d replicas built (%.1f s)" % (len(engines), time.perf_counter() - t0), file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    stats = run_replicas(engines, args.replica_moves, k=32)
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    moves = sum(st["generated"] for st in stats)
    sizes = [len(start["real"]) for start, _ in reps]
    out = {"workload": "%d independent single-element boxes, N log-uniform in [1000, 20000] (mean %.0f), density "
                       "0.05/A^3, PDC r[0,15]@0.05, translation 0.1 A, k<=32, own random streams per replica; "
                       "all replicas' windows of a step in one launch per kernel, walks and draws on the GPU"
                       % (len(reps), float(np.mean(sizes))),
           "moves_per_s": round(moves / el, 1), "moves": moves, "seconds": round(el, 3),
           "accepted_fraction": round(sum(st["accepted"] for st in stats) / moves, 4),
           "windows": sum(st["windows"] for st in stats)}
    print("bench: replicas done (%.1f moves/s)" % out["moves_per_s"], file=sys.stderr, flush=True)
    if not getattr(args, "replica_streams", True):
        return finish_replicas(args, out, reps, sizes)
    # the same replicas with one stream per engine (a window's kernels launched per replica)
    m2 = max(50, args.replica_moves // 4)
    t1 = time.perf_counter()
    stats2 = run_replicas(engines, m2, k=32, batched=False)
    torch.cuda.synchronize()
    el2 = time.perf_counter() - t1
    moves2 = sum(st["generated"] for st in stats2)
    out["per_engine_streams"] = {"moves_per_s": round(moves2 / el2, 1), "moves": moves2, "seconds": round(el2, 3)}
    print("bench: replicas on per-engine streams (%.1f moves/s)" % (moves2 / el2), file=sys.stderr, flush=True)
    return finish_replicas(args, out, reps, sizes)


def finish_replicas(args, out, reps, sizes):
    """The oracle engine on a sample of the replicas (cpu_baseline)."""
    if not args.no_cpu:
        # the oracle engine on four replicas spread over the size range, equal time each
        order = np.argsort(sizes)
        pick = [int(order[int(q * (len(order) - 1))]) for q in (0.125, 0.375, 0.625, 0.875)]
        cm = ct = 0.0
        for r in pick:
            start, pdc = reps[r]
            oe = cpu_engine_pdc(start, pdc)
            rate, n, el_r = cpu_moves_per_s(oe, args.cpu_seconds / 8)
            cm += n
            ct += el_r
        out["cpu_baseline"] = {"value": round(cm / ct, 2), "unit": "moves/s", "cores": 1, "kind": "port",
                               "sample": "%d moves on 4 replicas at the 12.5/37.5/62.5/87.5 %% size quantiles "
                                         "(oracle, 1 thread, equal time each) in %.1f s" % (cm, ct)}
        out["speedup_vs_cpu"] = round(out["moves_per_s"] / (cm / ct), 1)
    return out


def cpu_engine_pdc(start, pdc, seed=7):
    from oracle import chain
    from oracle import engine as oeng
    from fullrmc_amd.Constraints.PairDistributionConstraints import ATOMIC_NUMBER
    random.seed(seed)
    np.random.seed(seed)
    osys = oeng.System(start["real"], start["elements"], start["molecules"], start["basis"], start["recip"])
    w = chain.normalized_weighting(osys.counts, {e: ATOMIC_NUMBER[e] for e in osys.elements})
    oe = oeng.Engine(osys, [[i] for i in range(len(start["real"]))], oeng.Translation(0.1), [oeng.PDC(osys, pdc, w)])
    oe.initialize()
    return oe


def molecular_section(args):
    """500 rigid 13-atom rings (6,500 atoms, C/H/O) in a triclinic cell
    (42, 44, 46 A; 80, 95, 105 deg); PDC r[0,15]@0.05; whole-molecule moves
    of translation 0.2 A then rotation up to 10 deg, drawn by the Python
    generators (the reference's own NumPy calls) and evaluated on the
    device; beside the oracle engine on the same workload."""
    import torch
    from fullrmc_amd import systems
    from fullrmc_amd.Constraints import PairDistributionConstraint
    from fullrmc_amd.Engine import Engine
    from fullrmc_amd.Generators import MoveGeneratorCombinator, RotationGenerator, TranslationGenerator
    from oracle import chain
    from oracle import engine as oeng
    from fullrmc_amd.Constraints.PairDistributionConstraints import ATOMIC_NUMBER
    start = systems.ring_box(nmol=500, cell=(42.0, 44.0, 46.0, 80.0, 95.0, 105.0), seed=2050, min_centre=4.5)
    target = systems.jitter(start, 0.05, 2051)
    elem, pdc, _ = target_curves(start, target, 0.05, 300, np.array([1.0]))
    mols = {}
    for i, m in enumerate(start["molecules"]):
        mols.setdefault(int(m), []).append(i)
    groups = list(mols.values())
    random.seed(5)
    np.random.seed(5)
    E = Engine()
    E.set_system(start["real"], start["elements"], start["molecules"], start["basis"], start["recip"])
    E.set_groups(groups, MoveGeneratorCombinator([TranslationGenerator(0.2), RotationGenerator(10)]))
    E.add_constraints([PairDistributionConstraint(pdc)])
    E.run_device(50, k=16, pull=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = E.run_device(args.molecular_moves, k=16, pull=False)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    out = {"workload": "500 rigid 13-atom rings (6500 atoms, C/H/O), triclinic 42/44/46 A 80/95/105 deg, PDC "
                       "r[0,15]@0.05, molecule translation 0.2 A + rotation 10 deg (Python generators), k<=16",
           "moves_per_s": round(args.molecular_moves / el, 1), "moves": args.molecular_moves,
           "accepted_fraction": round(st["accepted"] / args.molecular_moves, 4), "windows": st["windows"]}
    out["distance_veto"] = distance_veto_record(args, start, groups, pdc)
    if not args.no_cpu:
        random.seed(5)
        np.random.seed(5)
        osys = oeng.System(start["real"], start["elements"], start["molecules"], start["basis"], start["recip"])
        w = chain.normalized_weighting(osys.counts, {e: ATOMIC_NUMBER[e] for e in osys.elements})
        oe = oeng.Engine(osys, groups, oeng.Combined([oeng.Translation(0.2), oeng.Rotation(10)]),
                         [oeng.PDC(osys, pdc, w)])
        oe.initialize()
        rate, n, el_c = cpu_moves_per_s(oe, args.cpu_seconds / 4)
        out["cpu_baseline"] = {"value": round(rate, 2), "unit": "moves/s", "cores": 1, "kind": "port",
                               "sample": "%d moves of the same workload (oracle, 1 thread) in %.1f s" % (n, el_c)}
        out["speedup_vs_cpu"] = round(out["moves_per_s"] / rate, 1)
    print("bench: molecular done (%.1f moves/s)" % out["moves_per_s"], file=sys.stderr, flush=True)
    return out


def distance_veto_record(args, start, groups, pdc):
    """The molecular workload with an inter-molecular distance constraint
    (2.4 A between atoms of different molecules, flexible): every proposal
    is vetoed or passed in Python (the constraint's own drop-in atomic
    distance calls) as its window is drawn, the PDC evaluated on the device;
    beside it the per-move drop-in loop on the same start."""
    import torch
    from fullrmc_amd.Constraints import InterMolecularDistanceConstraint, PairDistributionConstraint
    from fullrmc_amd.Engine import Engine
    from fullrmc_amd.Generators import MoveGeneratorCombinator, RotationGenerator, TranslationGenerator

    def engine():
        E = Engine()
        E.set_system(start["real"], start["elements"], start["molecules"], start["basis"], start["recip"])
        E.set_groups(groups, MoveGeneratorCombinator([TranslationGenerator(0.2), RotationGenerator(10)]))
        E.add_constraints([PairDistributionConstraint(pdc), InterMolecularDistanceConstraint(2.4)])
        return E

    n_dev = max(1, args.molecular_moves // 4)
    n_loop = max(1, args.molecular_moves // 10)
    random.seed(6)
    np.random.seed(6)
    E = engine()
    E.run_device(20, k=16, pull=False)
    torch.cuda.synchronize()
    g0, t0_ = E.generated, E.tried
    t0 = time.perf_counter()
    st = E.run_device(n_dev, k=16, pull=False)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    vetoed = (E.generated - g0) - (E.tried - t0_)
    random.seed(6)
    np.random.seed(6)
    P = engine()
    P.run(20)
    t0 = time.perf_counter()
    P.run(n_loop)
    el_p = time.perf_counter() - t0
    print("bench: distance veto done (%.1f moves/s)" % (n_dev / el), file=sys.stderr, flush=True)
    return {"workload": "the molecular workload plus InterMolecularDistanceConstraint(2.4 A, flexible): the veto "
                        "in Python per proposal as each window is drawn, the PDC on the device",
            "moves_per_s": round(n_dev / el, 1), "moves": n_dev, "vetoed_fraction": round(vetoed / n_dev, 4),
            "accepted_fraction": round(st["accepted"] / n_dev, 4), "windows": st["windows"],
            "dropin_per_move_loop_moves_per_s": round(n_loop / el_p, 1), "dropin_moves": n_loop}


def sq_transform_section(reps=50):
    """G(r) -> S(Q) for k proposals at once: the exact float32 sequential
    transform (bit-identical to the constraint) and the matrix-core
    contraction, with their error against a float64 transform; kernel_us is
    the median device time of the kernels (HIP events around them on the
    call stream), us_per_call the mean of back-to-back calls through the
    C ABI; the reference's NumPy expression per curve on the host beside
    them."""
    import torch
    from fullrmc_amd import _device as D
    from fullrmc_amd import _lib
    L = _lib.lib()
    out = {}
    for k, H, nQ, label in [(64, 1000, 500, "k64_H1000_Q500"), (128, 2000, 1000, "k128_H2000_Q1000")]:
        rng = np.random.default_rng(H)
        r = (0.005 + 0.01 * np.arange(H)).astype(np.float32)
        q = np.linspace(0.3, 30, nQ).astype(np.float32)
        Rs = r.reshape((-1, 1)) * (np.ones((H, 1), dtype=np.float32) * q)
        M = (r[1] - r[0]) * (np.sin(Rs) / q)
        G = (rng.standard_normal((k, H)) * np.exp(-r / 6)).astype(np.float32)
        d_g, d_m = D.upload(G), D.upload(M)
        d_s = D.empty((k, nQ), np.float32)
        S64 = G.astype(np.float64) @ M.astype(np.float64) + 1
        sec = {}
        for mode, mname in [(0, "exact"), (1, "mfma")]:
            def call():
                _lib.check(L.frmc_sq_transform(d_g.data_ptr(), k, H, d_m.data_ptr(), nQ, 1, mode, d_s.data_ptr(),
                                               D.stream()))
            call()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                call()
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1e3 / reps
            S = D.download(d_s)
            # device time of the transform kernels alone (HIP events on the call stream)
            L.frmc_set_timing(1)
            kus = []
            for _ in range(11):
                call()
                kus.append(1e3 * L.frmc_last_kernel_ms())
            L.frmc_set_timing(0)
            kus = sorted(kus)[5]
            tf = 2.0 * k * H * nQ / (kus * 1e-6) / 1e12
            sec[mname] = {"kernel_us": round(kus, 2), "us_per_call": round(us, 2), "tflops": round(tf, 3),
                          "max_abs_err_over_max_abs_S_vs_fp64": float("%.3g" % (np.max(np.abs(S - S64)) /
                                                                                 np.max(np.abs(S64))))}
        sec["mfma"]["roofline"] = {"bound": "mfma", "achieved": sec["mfma"]["tflops"], "peak": PEAK_FP32_TFLOPS,
                                   "unit": "TFLOP/s", "frac": round(sec["mfma"]["tflops"] / PEAK_FP32_TFLOPS, 4)}
        t0 = time.perf_counter()
        ncpu = 4
        for p in range(ncpu):
            np.sum(G[p].reshape((-1, 1)) * M, axis=0) + 1
        sec["cpu_numpy_us_per_curve"] = round((time.perf_counter() - t0) * 1e6 / ncpu, 1)
        sec["gpu_exact_us_per_curve"] = round(sec["exact"]["kernel_us"] / k, 3)
        out[label] = sec
    return out


def marker():
    """A short torch kernel on the current stream between synchronisations:
    in a rocprofv3 kernel trace its two launches bracket the timed region."""
    import torch
    torch.cuda.synchronize()
    torch.full((7,), 3, dtype=torch.int64, device="cuda").cumsum_(0)
    torch.cuda.synchronize()


def full_histogram_record(start, elem, dr, H, with_f64):
    """Full O(N^2) histogram: tuned and fixed-rule schedules (and the
    binary64-recompute mode), Gpairs/s on N(N-1)/2 pairs."""
    n = len(elem)
    gps, fms, _, sched = gpu_full_histogram(start, elem, dr, H)
    gps0, fms0, _, sched0 = gpu_full_histogram(start, elem, dr, H, autotune=False)
    # the roofline counts the pairs the kernel evaluated (a cell list skips
    # pairs in non-neighbouring cells); Gpairs/s stays on N(N-1)/2
    ev = sched.get("evaluated_pairs", -1)
    ev = ev if ev > 0 else n * (n - 1) // 2
    tf = ev / (fms * 1e-3) * OPS_PER_PAIR / 1e12
    out = {"gpairs_per_s": round(gps / 1e9, 2), "kernel_ms": round(fms, 3), "pairs": n * (n - 1) // 2,
           "evaluated_pairs": ev, "evaluated_gpairs_per_s": round(ev / (fms * 1e-3) / 1e9, 2),
           "schedule": sched,
           "fixed_rule": {"gpairs_per_s": round(gps0 / 1e9, 2), "kernel_ms": round(fms0, 3), "schedule": sched0},
           "roofline": {"bound": "valu", "compute": "fp32 VALU, 35 ops per evaluated pair", "achieved": round(tf, 3),
                        "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(tf / PEAK_FP32_TFLOPS, 4),
                        "attainable_peak": ATTAINABLE_FP32_TOPS,
                        "frac_of_attainable": round(tf / ATTAINABLE_FP32_TOPS, 4)}}
    if with_f64:
        gps64, fms64, re64, sched64 = gpu_full_histogram(start, elem, dr, H, float64=True)
        out["float64_recompute_mode"] = {"gpairs_per_s": round(gps64 / 1e9, 2), "kernel_ms": round(fms64, 3),
                                         "rechecked_pairs": re64, "schedule": sched64}
    return out


def engine_timed_run(args, E, n, shape, traffic_profile, steps, moves, k, log_first=None):
    """steps x moves generated moves through E (k speculative proposals per
    window) between two synchronisations, with HIP events around the
    kernels of every profile_every-th window; returns the counters, the
    per-kernel event times and the k_eval_a roofline."""
    import ctypes
    import torch
    from fullrmc_amd import _lib
    L = _lib.lib()
    h = E._device_engine(k)
    acc0 = E.accepted
    _lib.check(L.frmc_engine_profile(h, args.profile_every))
    marker()
    t0 = time.perf_counter()
    evaluated = windows = 0
    for i in range(steps):
        st = E.run_device(moves, k=k, pull=False, log=log_first if i == 0 else None)
        evaluated += st["evaluated"]
        windows += st["windows"]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    marker()
    nk = len(_lib.KERNEL_NAMES)
    ms = (ctypes.c_double * nk)()
    launches = (ctypes.c_int64 * nk)()
    _lib.check(L.frmc_engine_kernel_times(h, ms, launches))
    _lib.check(L.frmc_engine_profile(h, 0))
    kernels = {_lib.KERNEL_NAMES[i]: {"ms_total": round(ms[i], 3), "launches": int(launches[i]),
                                      "us_avg_with_events": round(1e3 * ms[i] / launches[i], 2)
                                      if launches[i] else None}
               for i in range(nk)}
    kmean = evaluated / max(1, windows)
    kwin = max(1, int(round(kmean)))
    # k_eval_a launched back to back at the run's mean window size, between
    # two HIP events on the engine stream: launch gaps and event markers
    # amortised, the figure a kernel trace's average duration compares with
    b2b = ctypes.c_float(0.0)
    marker()  # the third and fourth markers bracket these launches in a kernel trace
    _lib.check(L.frmc_engine_time_eval(h, kwin, 200, ctypes.byref(b2b)))
    marker()
    avg_s = b2b.value * 1e-3  # ms per launch
    # the tuned k_eval_a shape per window-size bucket, and every shape's time
    # at the mean window size
    shapes = (ctypes.c_int32 * 8)()
    _lib.check(L.frmc_engine_eval_shapes(h, shapes, 8))
    per_shape = {}
    try:
        for sh in list(range(0, 64, 4)) + [80, 84, 88, 92]:
            _lib.check(L.frmc_set_eval_shape(sh))
            t = ctypes.c_float(0.0)
            _lib.check(L.frmc_engine_time_eval(h, kwin, 50, ctypes.byref(t)))
            per_shape[str(sh)] = round(1e3 * t.value, 2)
    finally:
        _lib.check(L.frmc_set_eval_shape(args.eval_shape))
    traffic, tsrc = committed_traffic(traffic_profile, "k_eval_a", kwin)
    roof = kernel_roofline("k_eval_a", avg_s, n, kwin, shape, traffic, tsrc)
    if roof is not None:
        roof["duration_source"] = ("HIP events around 200 back-to-back launches at the run's mean window size "
                                   "(engine stream)")
        roof["in_run_us_with_events"] = kernels["k_eval_a"]["us_avg_with_events"]
        roof["tuned_shape_per_bucket"] = list(shapes)
        sh = shapes[min(7, max(0, (kwin - 1).bit_length()))]
        roof["shape"] = sh
        roof["kernel_instance"] = "k_eval_pc" if sh & 64 else "k_eval_a"
        roof["us_per_shape_at_mean_k"] = per_shape
    idel = _lib.KERNEL_NAMES.index("k_delta")
    roof_delta = kernel_roofline("k_delta", ms[idel] * 1e-3 / max(1, launches[idel]), n, kmean, shape)
    if roof_delta is not None:
        roof_delta["duration_source"] = "HIP events around the launches of every %d-th window" % args.profile_every
    return {"elapsed": elapsed, "moves": steps * moves, "accepted": E.accepted - acc0, "windows": windows,
            "evaluated": evaluated, "kernels": kernels, "roofline": roof, "roofline_k_delta": roof_delta}


def headline(args):
    """The 200,000-atom SiO2 box: full histograms, then the timed moves."""
    import torch
    t0 = time.perf_counter()
    start, elem, pdc, sfc = build_sio2_workload()
    n = len(elem)
    print("bench: SiO2 workload built (%d atoms, %.1f s)" % (n, time.perf_counter() - t0), file=sys.stderr,
          flush=True)
    fh = full_histogram_record(start, elem, 0.01, 2000, with_f64=True)
    print("bench: full histograms done (%.1f Gpairs/s)" % fh["gpairs_per_s"], file=sys.stderr, flush=True)
    E = make_engine(start, pdc, sfc, seed=3, dr=0.01)
    for _ in range(args.warmup):
        E.run_device(args.moves, k=args.k, pull=False)
    E._device_engine(args.k)  # created here when there is no warm-up
    snap = snapshot(E) if not args.no_cpu else None
    dlog = [] if snap is not None else None
    run = engine_timed_run(args, E, n, (2000, 2000, 1000, 3), "hbm_traffic_engine_sio2_200k.json", args.steps,
                           args.moves, args.k, log_first=dlog)
    value = run["moves"] / run["elapsed"]
    print("bench: timed region done: %.1f moves/s" % value, file=sys.stderr, flush=True)
    res = {"value": value, "elapsed": run["elapsed"], "n": n, "run": run, "full_histogram": fh}
    # plain form: one proposal per window, same box and constraints
    E1 = make_engine(start, pdc, sfc, seed=3, dr=0.01)
    E1.run_device(20, k=1, pull=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    E1.run_device(args.plain_moves, k=1, pull=False)
    torch.cuda.synchronize()
    plain = args.plain_moves / (time.perf_counter() - t1)
    E1._drop_device()
    res["plain"] = plain
    if snap is not None:
        # the oracle replays the first timed step from the same state and
        # random streams, for the CPU baseline and a parity check of the
        # timed trajectory
        oe = cpu_engine(start, pdc, sfc, 0.01, init_from=snap)
        random.setstate(snap["py"])
        np.random.set_state(snap["np"])
        olog = []
        tc = time.perf_counter()
        while time.perf_counter() - tc < args.cpu_seconds and len(olog) + 5 <= len(dlog):
            oe.run(5, olog)
        cel = time.perf_counter() - tc
        same = next((i for i, (a, b) in enumerate(zip(olog, dlog)) if a != b), len(olog))
        if same != len(olog):
            raise SystemExit("bench: the timed trajectory differs from the oracle at move %d: %r vs %r"
                             % (same, dlog[same], olog[same]))
        threads = len(os.sched_getaffinity(0))
        cgp, cpairs = cpu_full_histogram_sample(start, elem, 0.01, 2000, 200, threads)
        res["cpu"] = {"value": round(len(olog) / cel, 2), "unit": "moves/s", "cores": 1, "kind": "port",
                      "sample": "the first %d moves of the timed region replayed by the oracle (C restatement of "
                                "the reference loops + NumPy chain, 1 thread) from the same state and random "
                                "streams, in %.1f s" % (len(olog), cel),
                      "parity_checked_moves": len(olog),
                      "full_histogram_gpairs_per_s": round(cgp / 1e9, 4), "full_histogram_cores": threads,
                      "full_histogram_sample": "200 row atoms vs all later atoms (%.3g pairs), oracle C loop, "
                                               "OpenMP with private bins" % cpairs}
        res["cpu_full_hist_gps"] = cgp
    E._drop_device()
    return res


def swap_moves_record(start, pdc, sfc):
    """Swap moves on the same B2 box: the first element's atoms exchange
    positions with a randomly drawn atom of the second (SwapPositionsGenerator),
    the second element's atoms translate by 0.1 A.  The device engine in
    speculative windows of Python-drawn proposals (each naming its moved
    atoms), and the reference's loop over the drop-in calls."""
    import torch
    from fullrmc_amd.Generators import SwapPositionsGenerator, TranslationGenerator
    els = start["elements"]
    n = len(els)
    first, second = sorted(set(els))[:2]
    partners = [[j] for j in range(n) if els[j] == second]

    def engine():
        E = make_engine(start, pdc, sfc, seed=2)
        sw, tr = SwapPositionsGenerator(swapList=partners), TranslationGenerator(0.1)
        E.set_groups([[i] for i in range(n)], [sw if els[i] == first else tr for i in range(n)])
        return E

    E = engine()
    E.run_device(200, k=16, pull=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = E.run_device(2000, k=16, pull=False)
    torch.cuda.synchronize()
    dev = 2000 / (time.perf_counter() - t0)
    E._drop_device()
    E2 = engine()
    E2.run(2)
    t1 = time.perf_counter()
    E2.run(30)
    dropin = 30 / (time.perf_counter() - t1)
    return {"workload": "B2 alloy 13500 atoms: %s atoms swap positions with a drawn %s atom, %s atoms translate "
                        "0.1 A; PDC + SFC as above" % (first, second, second),
            "device_engine_moves_per_s": round(dev, 1), "windows": st["windows"],
            "dropin_per_move_loop_moves_per_s": round(dropin, 1)}


def removal_moves_record(start, pdc, sfc):
    """Removals on the same B2 box: 500 extra groups remove a drawn atom of
    the second element (AtomsRemoveGenerator, at most 20 in all) beside the
    translations of every atom.  The device engine (a removal evaluated on
    the host at the end of its window, an accepted one rebuilding the engine
    on the remaining atoms), and the reference's loop over the drop-in
    calls."""
    import torch
    from fullrmc_amd.Generators import AtomsRemoveGenerator, TranslationGenerator
    els = start["elements"]
    n = len(els)
    second = sorted(set(els))[1]
    pool = [j for j in range(n) if els[j] == second]

    def engine():
        E = make_engine(start, pdc, sfc, seed=3)
        rm, tr = AtomsRemoveGenerator(maximumCollected=20, atomsList=pool), TranslationGenerator(0.1)
        E.set_groups([[i] for i in range(n)] + [[0]] * 500, [tr] * n + [rm] * 500)
        return E

    E = engine()
    E.run_device(100, k=16, pull=False)
    torch.cuda.synchronize()
    r0 = E.removed[1]
    t0 = time.perf_counter()
    st = E.run_device(1000, k=16, pull=False)
    torch.cuda.synchronize()
    dev = 1000 / (time.perf_counter() - t0)
    removed = E.removed[1] - r0
    E._drop_device()
    E2 = engine()
    E2.run(2)
    t1 = time.perf_counter()
    E2.run(30)
    dropin = 30 / (time.perf_counter() - t1)
    return {"workload": "B2 alloy 13500 atoms: translations of 0.1 A, and 500 groups removing a drawn %s atom "
                        "(at most 20); PDC + SFC as above" % second,
            "device_engine_moves_per_s": round(dev, 1), "windows": st["windows"], "removed_in_timed_moves": removed,
            "dropin_per_move_loop_moves_per_s": round(dropin, 1)}


def b2_section(args):
    """The 13,500-atom B2 alloy: engine moves (k <= 64), plain form, the
    reference's loop over the drop-in calls, full histogram, CPU oracle."""
    import torch
    start, elem, pdc, sfc = build_workload()
    n = len(elem)
    E = make_engine(start, pdc, sfc, seed=1)
    E.run_device(args.b2_moves, k=64, pull=False)
    run = engine_timed_run(args, E, n, (1000, 1000, 500, 3), "hbm_traffic_engine_b2_13500.json", 5,
                           args.b2_moves, 64)
    out = {"workload": "B2 alloy 13500 atoms (15x15x30 cells, a = 3.01 A), PDC r[0,20]@0.02 + SFC Q[0.5,25]x500, "
                       "single-atom translation 0.1 A, k<=64",
           "moves_per_s": round(run["moves"] / run["elapsed"], 1), "moves": run["moves"],
           "accepted_fraction": round(run["accepted"] / run["moves"], 4), "windows": run["windows"],
           "proposals_evaluated": run["evaluated"], "roofline_k_eval_a": run["roofline"],
           "full_histogram": full_histogram_record(start, elem, 0.02, 1000, with_f64=False)}
    E._drop_device()
    E1 = make_engine(start, pdc, sfc, seed=1)
    E1.run_device(50, k=1, pull=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    E1.run_device(1000, k=1, pull=False)
    torch.cuda.synchronize()
    plain = 1000 / (time.perf_counter() - t1)
    E1._drop_device()
    out["plain_one_move_per_window_moves_per_s"] = round(plain, 1)
    # algorithmic bytes of one move's delta: 20 B per other atom (xyz, molecule, element)
    out["plain_one_move_per_window_gb_per_s"] = round(20.0 * n * plain / 1e9, 3)
    E2 = make_engine(start, pdc, sfc, seed=1)
    E2.run(2)
    t2 = time.perf_counter()
    E2.run(30)
    out["dropin_per_move_loop_moves_per_s"] = round(30 / (time.perf_counter() - t2), 1)
    E2._drop_device()
    out["swap_moves"] = swap_moves_record(start, pdc, sfc)
    out["removal_moves"] = removal_moves_record(start, pdc, sfc)
    if not args.no_cpu:
        oe = cpu_engine(start, pdc, sfc, 0.02)
        cps, cn, cel = cpu_moves_per_s(oe, args.cpu_seconds / 2)
        out["cpu_baseline"] = {"value": round(cps, 2), "unit": "moves/s", "cores": 1, "kind": "port",
                               "sample": "%d moves of the same workload (oracle, 1 thread) in %.1f s" % (cn, cel)}
        out["speedup_vs_cpu"] = round(out["moves_per_s"] / cps, 1)
    print("bench: B2 done (%.1f moves/s)" % out["moves_per_s"], file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--moves", type=int, default=5000, help="generated moves per step")
    ap.add_argument("--k", type=int, default=128, help="speculative proposals per window (at most)")
    ap.add_argument("--plain-moves", type=int, default=1000, help="moves of the one-proposal-per-window form")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline (no B2, transform, replica or molecular sections)")
    # every 32nd window: about 100 timed launches per kernel per 3,000
    # windows; every 8th cost 4 % of the timed region's moves/s
    ap.add_argument("--profile-every", type=int, default=32,
                    help="time the kernels of every n-th window with HIP events (0: off)")
    ap.add_argument("--b2-moves", type=int, default=4000, help="moves per timed step of the B2 section (0: skip)")
    ap.add_argument("--replicas", type=int, default=256, help="independent replicas run together (0: skip)")
    ap.add_argument("--replica-moves", type=int, default=3000, help="generated moves per replica")
    # off by default: under a rocprofv3 kernel trace this form (256 engine
    # streams fed by 16 host threads) has crashed in the runtime's launch path
    ap.add_argument("--replica-streams", action=argparse.BooleanOptionalAction, default=False,
                    help="also time the replicas with one stream per engine")
    ap.add_argument("--molecular-moves", type=int, default=1000,
                    help="moves of the 6,500-atom molecular triclinic box (0: skip)")
    ap.add_argument("--eval-shape", type=int, default=-1,
                    help="force one window-evaluation shape (0..60 step 4; -1: tuned per window size)")
    ap.add_argument("--window-loop", type=int, default=-1, choices=(-1, 0, 1),
                    help="where windows are walked and drawn: -1 default (host for a single engine, device for "
                         "batched replicas), 0 host, 1 device")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB device-to-device copy (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the pair path runs on a single MI355X; "
              "a sharded multi-GPU run is not implemented.", file=sys.stderr)
        sys.exit(2)
    extras = not args.no_extras

    replica_boxes = None
    if extras and args.replicas > 0:
        t0 = time.perf_counter()
        replica_boxes = build_replica_boxes(args.replicas)  # before the GPU is initialised (fork pool)
        print("bench: %d replica boxes built (%.1f s)" % (len(replica_boxes), time.perf_counter() - t0),
              file=sys.stderr, flush=True)

    import torch
    from fullrmc_amd import _lib
    from fullrmc_amd._device import require_gpu
    require_gpu()
    _lib.check(_lib.lib().frmc_set_eval_shape(args.eval_shape))
    _lib.check(_lib.lib().frmc_set_window_loop(args.window_loop))

    if args.copy_calibration:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.int32, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    H = headline(args)
    run = H["run"]
    moves = run["moves"]
    result = {"metric": "bit-exact-trajectory RMC moves/s + full-histogram Gpairs/s, GPU vs host CPU",
              "value": round(H["value"], 1), "unit": "moves/s", "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": round(1e3 * H["elapsed"] / args.steps, 3),
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
              "data": "synthetic",
              "config": {"workload": "SiO2-like 200000 atoms (66667 Si + 133333 O), cube 145 A, min distance "
                                     "1.5 A; PDC r[0,20]@0.01 (2000 bins) + SFC Q[0.3,30]x1000 (own r grid "
                                     "[0,20]@0.01), single-atom translation 0.1 A, k<=%d speculative windows; "
                                     "%d warm-up moves (pre-equilibration), %d timed moves"
                                     % (args.k, args.warmup * args.moves, moves),
                         "atoms": H["n"], "moves_per_step": args.moves, "k": args.k},
              "roofline": run["roofline"],
              "roofline_k_delta": run["roofline_k_delta"],
              "accepted_fraction": round(run["accepted"] / moves, 4),
              "windows": run["windows"], "proposals_evaluated": run["evaluated"], "kernels": run["kernels"],
              "full_histogram": H["full_histogram"],
              "plain_one_move_per_window_moves_per_s": round(H["plain"], 1),
              "plain_one_move_per_window_gb_per_s": round(20.0 * H["n"] * H["plain"] / 1e9, 3)}
    if "cpu" in H:
        result["cpu_baseline"] = H["cpu"]
        result["speedup_vs_cpu"] = round(H["value"] / H["cpu"]["value"], 1)
        result["full_histogram_speedup_vs_cpu_all_cores"] = round(
            1e9 * H["full_histogram"]["gpairs_per_s"] / H["cpu_full_hist_gps"], 1)
    if extras:
        if args.b2_moves > 0:
            result["b2_13500"] = b2_section(args)
        result["sq_transform"] = sq_transform_section()
        if args.replicas > 0:
            result["replicas"] = replicas_section(args, replica_boxes)
        if args.molecular_moves > 0:
            result["molecular_6500"] = molecular_section(args)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
