"""Benchmark of the pair-distance / pair-histogram hot path on one MI355X.

Workload (N=1): the binary B2 alloy of 13,500 atoms (15x15x30 cells,
a = 3.01 A, 50/50 species, Gaussian jitter sigma = 0.08 A) with a
PairDistributionConstraint (r in [0,20] A at 0.02 A) and a
StructureFactorConstraint (own grid r in [0,20] A at 0.02 A, Q = 0.5..25 1/A,
500 points), single-atom translation moves of amplitude 0.1 A.  The targets
are float64 G(r)/S(Q) of an independently jittered copy of the lattice.

One step = ``--moves`` generated RMC moves through the device-resident
engine with k speculative proposals per window (bit-exact trajectory).
``value`` = generated moves per second over the timed steps.  Also reported:
the plain one-proposal-per-window form and the drop-in per-move loop on the
same GPU, full-histogram Gpairs/s (one O(N^2) recompute), the roofline of the
dominant kernel from HIP events recorded around every launch during the
timed region, and the CPU oracle timed on this host.
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
    L.frmc_set_timing(0)
    sched = tuning.last_full_schedule()
    tuning.set_autotune(True)
    ms = sorted(ms[1:])[len(ms[1:]) // 2]
    pairs = n * (n - 1) / 2
    return pairs / (ms * 1e-3), ms, int(rechecked.value), sched


def cpu_engine(start, pdc, sfc, dr, seed=7, init_from=None):
    """The oracle engine on the workload; init_from (a device Engine, synced to
    the host) supplies the coordinates and slot-exact histograms instead of an
    O(N^2) oracle recompute."""
    from oracle import chain
    from oracle import engine as oeng
    from fullrmc_amd.Constraints.PairDistributionConstraints import ATOMIC_NUMBER
    random.seed(seed)
    np.random.seed(seed)
    real = start["real"] if init_from is None else init_from.realCoordinates
    osys = oeng.System(real, start["elements"], start["molecules"], start["basis"], start["recip"])
    w = chain.normalized_weighting(osys.counts, {e: ATOMIC_NUMBER[e] for e in osys.elements})
    oe = oeng.Engine(osys, [[i] for i in range(len(start["real"]))], oeng.Translation(0.1),
                     [oeng.PDC(osys, pdc, w), oeng.SFC(osys, sfc, w, 0.0, 20.0, dr)])
    if init_from is not None:
        assert osys.box.tobytes() == np.asarray(init_from.boxCoordinates, np.float32).tobytes()
    if init_from is None:
        oe.initialize()
    else:
        for c, dc in zip(oe.constraints, init_from.constraints):
            c.intra = np.array(dc.data["intra"], dtype=np.float32)
            c.inter = np.array(dc.data["inter"], dtype=np.float32)
            c.sd = c.chi2(c.intra, c.inter)
        oe.tse = chain.total_error([c.sd for c in oe.constraints])
    return oe


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


def eval_flops_per_proposal(H_pdc=1000, H_sfc=1000, nQ=500, T=3):
    """k_eval per proposal: G(r) (T multiply-divide-add + divide, subtract,
    multiply per bin), S(Q) (one product and one add per (r, Q)), squared
    deviations (3 per point) and the pairwise sums (1 per point)."""
    gr = (3 * T + 3) * (H_pdc + H_sfc)
    return gr + 2.0 * H_sfc * nQ + 4.0 * (H_pdc + nQ)


def kernel_roofline(name, avg_s, n, props_per_launch):
    """Algorithmic work of one launch of the named engine kernel, against the
    FP32 vector peak (the pair path issues no matrix instructions)."""
    if name == "k_delta":
        flops = OPS_PER_PAIR * 2.0 * n * props_per_launch      # old + new position vs all atoms
    elif name == "k_eval_a":
        flops = eval_flops_per_proposal() * props_per_launch
    elif name == "k_eval_b":
        flops = 2.0 * 500 * props_per_launch                   # SFC squared deviations summed, totals
    elif name == "k_commit":
        flops = OPS_PER_PAIR * 2.0 * n
    else:
        return None
    if not avg_s > 0:
        return None
    tf = flops / avg_s / 1e12
    # HBM bytes per launch: the PMC per-proposal figure times this timed
    # region's proposals per launch
    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic_engine_b2_13500.json")) as f:
            traffic = int(round(json.load(f)["kernels"][name]["traffic_bytes_per_proposal"] * props_per_launch))
    except (OSError, KeyError, ValueError):
        pass
    return {"bound": "mfma", "compute": "fp32 VALU (no matrix instructions; MI355X fp32 vector and dense "
            "matrix peaks are both 157.3 TFLOP/s)", "achieved": round(tf, 4), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tf / PEAK_FP32_TFLOPS, 6), "traffic": traffic,
            "traffic_source": "profiles/hbm_traffic_engine_b2_13500.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                              "copy-calibrated, per proposal x proposals per launch)" if traffic is not None else None,
            "kernel": name, "avg_launch_us": round(avg_s * 1e6, 2), "flops_per_launch": flops}


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
                       "all replicas' windows of a step in one launch per kernel"
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


def sio2_section(args):
    """200,000-atom SiO2 box: full histogram (float32 reference binning and
    the binary64-recompute mode) and device moves with k <= 128, beside the
    CPU oracle on bounded samples."""
    import torch
    t0 = time.perf_counter()
    start, elem, pdc, sfc = build_sio2_workload()
    n = len(elem)
    print("bench: SiO2 workload built (%d atoms, %.1f s)" % (n, time.perf_counter() - t0), file=sys.stderr,
          flush=True)
    out = {"workload": "SiO2-like 200000 atoms (66667 Si), cube 145 A, min distance 1.5 A; PDC r[0,20]@0.01 "
                       "(2000 bins) + SFC Q[0.3,30]x1000, single-atom translation 0.1 A, k<=128"}
    gps, fms, _, sched = gpu_full_histogram(start, elem, 0.01, 2000)
    gps0, fms0, _, sched0 = gpu_full_histogram(start, elem, 0.01, 2000, autotune=False)
    gps64, fms64, re64, sched64 = gpu_full_histogram(start, elem, 0.01, 2000, float64=True)
    tf = gps * OPS_PER_PAIR / 1e12
    out["full_histogram"] = {"gpairs_per_s": round(gps / 1e9, 2), "kernel_ms": round(fms, 3),
                             "pairs": n * (n - 1) // 2, "schedule": sched,
                             "fixed_rule": {"gpairs_per_s": round(gps0 / 1e9, 2), "kernel_ms": round(fms0, 3),
                                            "schedule": sched0},
                             "roofline": {"bound": "mfma", "compute": "fp32 VALU", "achieved": round(tf, 3),
                                          "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                          "frac": round(tf / PEAK_FP32_TFLOPS, 4)},
                             "float64_recompute_mode": {"gpairs_per_s": round(gps64 / 1e9, 2),
                                                        "kernel_ms": round(fms64, 3), "rechecked_pairs": re64,
                                                        "schedule": sched64}}
    E = make_engine(start, pdc, sfc, seed=3, dr=0.01)
    E.run_device(500, k=128, pull=False)
    torch.cuda.synchronize()
    acc0 = E.accepted
    t1 = time.perf_counter()
    st = E.run_device(args.sio2_moves, k=128, pull=False)
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    out["moves_per_s"] = round(args.sio2_moves / el, 1)
    out["moves"] = args.sio2_moves
    out["accepted_fraction"] = round((E.accepted - acc0) / args.sio2_moves, 4)
    out["windows"] = st["windows"]
    out["proposals_evaluated"] = st["evaluated"]
    print("bench: SiO2 moves done (%.1f moves/s)" % out["moves_per_s"], file=sys.stderr, flush=True)
    if not args.no_cpu:
        E.sync_from_device()
        oe = cpu_engine(start, pdc, sfc, 0.01, init_from=E)
        cps, cn, cel = cpu_moves_per_s(oe, args.cpu_seconds / 2)
        threads = len(os.sched_getaffinity(0))
        cgp, cpairs = cpu_full_histogram_sample(start, elem, 0.01, 2000, 200, threads)
        out["cpu_baseline"] = {"value": round(cps, 2), "unit": "moves/s", "cores": 1, "kind": "port",
                               "sample": "%d moves (oracle, 1 thread) in %.1f s" % (cn, cel),
                               "full_histogram_gpairs_per_s": round(cgp / 1e9, 4), "full_histogram_cores": threads,
                               "full_histogram_sample": "200 row atoms vs all later atoms (%.3g pairs)" % cpairs}
        out["speedup_vs_cpu"] = round(out["moves_per_s"] / cps, 1)
        out["full_histogram_speedup_vs_cpu_all_cores"] = round(gps / cgp, 1)
    return out


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


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--moves", type=int, default=4000, help="generated moves per step")
    ap.add_argument("--k", type=int, default=64, help="speculative proposals per window")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    # every 32nd window: about 100 timed launches per kernel; every 8th cost
    # 4 % of the timed region's moves/s in event records
    ap.add_argument("--profile-every", type=int, default=32,
                    help="time the kernels of every n-th window with HIP events (0: off)")
    ap.add_argument("--sio2", action=argparse.BooleanOptionalAction, default=True,
                    help="also run the 200,000-atom SiO2 box (full histograms and k=128 moves)")
    ap.add_argument("--sio2-moves", type=int, default=3000)
    ap.add_argument("--replicas", type=int, default=256, help="independent replicas run together (0: skip)")
    ap.add_argument("--replica-moves", type=int, default=3000, help="generated moves per replica")
    ap.add_argument("--replica-streams", action=argparse.BooleanOptionalAction, default=True,
                    help="also time the replicas with one stream per engine")
    ap.add_argument("--molecular-moves", type=int, default=1000,
                    help="moves of the 6,500-atom molecular triclinic box (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB device-to-device copy (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the pair path runs on a single MI355X; "
              "a sharded multi-GPU run is not implemented.", file=sys.stderr)
        sys.exit(2)

    replica_boxes = None
    if args.replicas > 0:
        t0 = time.perf_counter()
        replica_boxes = build_replica_boxes(args.replicas)  # before the GPU is initialised (fork pool)
        print("bench: %d replica boxes built (%.1f s)" % (len(replica_boxes), time.perf_counter() - t0),
              file=sys.stderr, flush=True)

    import torch
    from fullrmc_amd import _lib
    from fullrmc_amd._device import require_gpu
    require_gpu()
    L = _lib.lib()

    if args.copy_calibration:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.int32, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    start, elem, pdc, sfc = build_workload()
    n = len(elem)
    print("bench: workload built (%d atoms)" % n, file=sys.stderr, flush=True)
    E = make_engine(start, pdc, sfc, seed=1)
    for _ in range(args.warmup):
        E.run_device(args.moves, k=args.k, pull=False)
    h = E._device_engine(args.k)  # created here when there is no warm-up
    acc0 = E.accepted
    _lib.check(L.frmc_engine_profile(h, args.profile_every))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evaluated = windows = 0
    for _ in range(args.steps):
        st = E.run_device(args.moves, k=args.k, pull=False)
        evaluated += st["evaluated"]
        windows += st["windows"]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    nk = len(_lib.KERNEL_NAMES)
    ms = (ctypes.c_double * nk)()
    launches = (ctypes.c_int64 * nk)()
    _lib.check(L.frmc_engine_kernel_times(h, ms, launches))
    _lib.check(L.frmc_engine_profile(h, 0))
    moves = args.steps * args.moves
    value = moves / elapsed
    accepted = E.accepted - acc0
    print("bench: timed region done: %.1f moves/s" % value, file=sys.stderr, flush=True)
    kernels = {_lib.KERNEL_NAMES[i]: {"ms_total": round(ms[i], 3), "launches": int(launches[i]),
                                      "us_avg": round(1e3 * ms[i] / launches[i], 2) if launches[i] else None}
               for i in range(nk)}
    # the roofline kernel: k_eval_a (G(r), S(Q) and chi-square of every proposal)
    dom = _lib.KERNEL_NAMES.index("k_eval_a")
    name = _lib.KERNEL_NAMES[dom]
    avg_s = ms[dom] * 1e-3 / max(1, launches[dom])
    roof = kernel_roofline(name, avg_s, n, evaluated / max(1, windows))
    # the same kernel launched back to back (launch gaps amortised): the figure
    # a rocprofv3 kernel trace's average duration compares with
    kwin = max(1, int(round(evaluated / max(1, windows))))
    b2b = ctypes.c_float(0.0)
    _lib.check(L.frmc_engine_time_eval(h, kwin, 200, ctypes.byref(b2b)))
    if roof is not None:
        roof["back_to_back_us"] = round(1e3 * b2b.value, 2)
        roof["back_to_back_k"] = kwin
    # the batched delta kernel against the same FP32 peak (70 ops per
    # proposal and other atom, old and new position)
    idel = _lib.KERNEL_NAMES.index("k_delta")
    roof_delta = kernel_roofline("k_delta", ms[idel] * 1e-3 / max(1, launches[idel]), n,
                                 evaluated / max(1, windows))

    result = {"metric": "bit-exact-trajectory RMC moves/s + full-histogram Gpairs/s, GPU vs host CPU",
              "value": round(value, 1), "unit": "moves/s", "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3),
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
              "data": "synthetic",
              "config": {"workload": "B2 alloy 13500 atoms, PDC r[0,20]@0.02 + SFC Q[0.5,25]x500, "
                                     "single-atom translation 0.1 A, k=%d speculative windows" % args.k,
                         "atoms": n, "moves_per_step": args.moves, "k": args.k},
              "roofline": roof,
              "roofline_k_delta": roof_delta,
              "accepted_fraction": round(accepted / moves, 4),
              "windows": windows, "proposals_evaluated": evaluated, "kernels": kernels}

    if not args.no_extras:
        # full-histogram recompute (one O(N^2) pass on the PDC grid)
        gps, fms, _, sched = gpu_full_histogram(start, elem, 0.02, 1000)
        gps0, fms0, _, sched0 = gpu_full_histogram(start, elem, 0.02, 1000, autotune=False)
        tf = gps * OPS_PER_PAIR / 1e12
        result["full_histogram"] = {"gpairs_per_s": round(gps / 1e9, 3), "kernel_ms": round(fms, 3),
                                    "pairs": n * (n - 1) // 2, "schedule": sched,
                                    "fixed_rule": {"gpairs_per_s": round(gps0 / 1e9, 3), "kernel_ms": round(fms0, 3),
                                                   "schedule": sched0},
                                    "roofline": {"bound": "mfma", "compute": "fp32 VALU", "achieved": round(tf, 3),
                                                 "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                                 "frac": round(tf / PEAK_FP32_TFLOPS, 4)}}
        # plain form: one proposal per window
        E1 = make_engine(start, pdc, sfc, seed=1)
        E1.run_device(50, k=1, pull=False)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        E1.run_device(1000, k=1, pull=False)
        torch.cuda.synchronize()
        plain = 1000 / (time.perf_counter() - t1)
        result["plain_one_move_per_window_moves_per_s"] = round(plain, 1)
        # algorithmic bytes of one move's delta: 20 B per other atom (xyz,
        # molecule, element)
        result["plain_one_move_per_window_gb_per_s"] = round(20.0 * n * plain / 1e9, 3)
        # the reference's loop structure calling the drop-in extensions per move
        E2 = make_engine(start, pdc, sfc, seed=1)
        E2.run(2)
        t2 = time.perf_counter()
        E2.run(30)
        result["dropin_per_move_loop_moves_per_s"] = round(30 / (time.perf_counter() - t2), 1)
        result["sq_transform"] = sq_transform_section()
        print("bench: extras done", file=sys.stderr, flush=True)

    if args.sio2:
        result["sio2_200k"] = sio2_section(args)
    if args.replicas > 0:
        result["replicas"] = replicas_section(args, replica_boxes)
    if args.molecular_moves > 0:
        result["molecular_6500"] = molecular_section(args)

    if not args.no_cpu:
        oe = cpu_engine(start, pdc, sfc, 0.02)
        cps, cn, cel = cpu_moves_per_s(oe, args.cpu_seconds)
        threads = len(os.sched_getaffinity(0))
        # the same oracle loop with its per-move histograms on every core
        oe.s.ncores = threads
        aps, an, ael = cpu_moves_per_s(oe, args.cpu_seconds / 3)
        cgp, cpairs = cpu_full_histogram_sample(start, elem, 0.02, 1000, 400, threads)
        result["cpu_baseline"] = {"value": round(cps, 2), "unit": "moves/s", "cores": 1, "kind": "port",
                                  "sample": "%d moves of the same workload (oracle: C restatement of the reference "
                                            "loops + NumPy chain, 1 thread) in %.1f s" % (cn, cel),
                                  "all_cores_value": round(aps, 2), "all_cores": threads,
                                  "all_cores_sample": "%d moves, per-move histogram loops on %d threads (OpenMP, "
                                                      "private bins) in %.1f s" % (an, threads, ael),
                                  "full_histogram_gpairs_per_s": round(cgp / 1e9, 4),
                                  "full_histogram_cores": threads,
                                  "full_histogram_sample": "%d row atoms vs all later atoms (%.3g pairs), oracle C "
                                                           "loop, OpenMP with private bins" % (400, cpairs)}
        result["speedup_vs_cpu"] = round(value / cps, 1)
        result["speedup_vs_cpu_all_cores"] = round(value / aps, 1)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
