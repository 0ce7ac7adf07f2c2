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


def build_workload(seed=2026):
    from fullrmc_amd import systems
    from fullrmc_amd.Constraints.PairDistributionConstraints import ATOMIC_NUMBER, normalized_weighting
    from fullrmc_amd.Core.pairs_histograms import full_pairs_histograms_coords

    start = systems.b2_alloy(cells=(15, 15, 30), a=3.01, sigma=0.08, seed=seed)
    target = systems.b2_alloy(cells=(15, 15, 30), a=3.01, sigma=0.08, seed=seed + 1)
    els = sorted(set(start["elements"]))
    lut = {e: i for i, e in enumerate(els)}
    elem = np.array([lut[e] for e in start["elements"]], dtype=np.int32)
    counts = {e: start["elements"].count(e) for e in els}
    n = len(elem)
    V = float(systems.volume(start["basis"]))
    rho = n / V
    w = {k: float(v) for k, v in normalized_weighting(counts, {e: ATOMIC_NUMBER[e] for e in els}).items()}
    # float64 target G(r) on r = 0.01 + 0.02 k from the target box's pair counts
    tbox = systems.box_coordinates(target["recip"], target["real"])
    hi, he = full_pairs_histograms_coords(tbox, target["basis"], True, target["molecules"], elem, 2,
                                          np.float32(0.0), np.float32(20.0), np.float32(0.02), 1000)
    r = 0.01 + 0.02 * np.arange(1000)
    edges = 0.02 * np.arange(1001)
    shell = 4.0 / 3.0 * np.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    g = np.zeros(1000)
    for a in range(2):
        for b in range(a, 2):
            ea, eb = els[a], els[b]
            if a == b:
                nij = hi[a, a].astype(np.float64) + he[a, a]
                D = counts[ea] * (counts[ea] - 1) / 2.0 / V
            else:
                nij = (hi[a, b].astype(np.float64) + hi[b, a] + he[a, b] + he[b, a])
                D = counts[ea] * counts[eb] / V
            g += w[ea + "-" + eb] * nij / D / shell
    G = 4 * np.pi * r * rho * (g - 1)
    q = np.linspace(0.5, 25.0, 500)
    S = 1 + np.sum(G[:, None] * 0.02 * np.sin(np.outer(r, q)) / q, axis=0)
    pdc = np.stack([r, G], 1).astype(np.float32)
    sfc = np.stack([q, S], 1).astype(np.float32)
    return start, elem, pdc, sfc


def make_engine(start, pdc, sfc, seed):
    from fullrmc_amd.Constraints import PairDistributionConstraint, StructureFactorConstraint
    from fullrmc_amd.Engine import Engine
    from fullrmc_amd.Generators import TranslationGenerator
    random.seed(seed)
    np.random.seed(seed)
    E = Engine()
    E.set_system(start["real"], start["elements"], start["molecules"], start["basis"], start["recip"])
    E.set_move_generator(TranslationGenerator(0.1))
    E.add_constraints([PairDistributionConstraint(pdc),
                       StructureFactorConstraint(sfc, rmin=0.0, rmax=20.0, dr=0.02)])
    return E


def gpu_full_histogram(start, elem, reps=3):
    """One O(N^2) recompute on the PDC grid: kernel time by HIP events."""
    from fullrmc_amd import _device as D
    from fullrmc_amd import _lib, systems
    L = _lib.lib()
    box = systems.box_coordinates(start["recip"], start["real"])
    n = len(box)
    d_c, d_b = D.upload(box), D.upload(start["basis"])
    d_m, d_e = D.upload(start["molecules"]), D.upload(elem)
    counts = D.empty((2 * 2 * 2 * 1000 + 1,), np.int32)
    L.frmc_set_timing(1)
    ms = []
    for _ in range(reps + 1):
        _lib.check(L.frmc_full_pairs_counts(d_c.data_ptr(), n, d_b.data_ptr(), 1, d_m.data_ptr(), d_e.data_ptr(),
                                            2, 0.0, 20.0, 0.02, 1000, counts.data_ptr(), D.stream()))
        ms.append(L.frmc_last_kernel_ms())
    L.frmc_set_timing(0)
    ms = sorted(ms[1:])[len(ms[1:]) // 2]
    pairs = n * (n - 1) / 2
    return pairs / (ms * 1e-3), ms


def cpu_baseline(start, pdc, sfc, seconds):
    """The CPU oracle (C restatement of the reference loops + the NumPy chain),
    one thread, on a bounded sample of the same workload."""
    from oracle import chain
    from oracle import engine as oeng
    from fullrmc_amd.Constraints.PairDistributionConstraints import ATOMIC_NUMBER
    random.seed(7)
    np.random.seed(7)
    osys = oeng.System(start["real"], start["elements"], start["molecules"], start["basis"], start["recip"])
    w = chain.normalized_weighting(osys.counts, {e: ATOMIC_NUMBER[e] for e in osys.elements})
    oe = oeng.Engine(osys, [[i] for i in range(len(start["real"]))], oeng.Translation(0.1),
                     [oeng.PDC(osys, pdc, w), oeng.SFC(osys, sfc, w, 0.0, 20.0, 0.02)])
    t0 = time.perf_counter()
    oe.initialize()
    t_full = time.perf_counter() - t0
    n = 0
    t0 = time.perf_counter()
    while True:
        oe.run(5)
        n += 5
        el = time.perf_counter() - t0
        if el > seconds:
            break
    npairs = len(start["real"]) * (len(start["real"]) - 1) / 2
    return n / el, n, el, 2 * npairs / t_full


def eval_flops_per_proposal(H_pdc=1000, H_sfc=1000, nQ=500, T=3):
    """k_eval per proposal: G(r) (T multiply-divide-add + divide, subtract,
    multiply per bin), S(Q) (one product and one add per (r, Q)), squared
    deviations (3 per point) and the pairwise sums (1 per point)."""
    gr = (3 * T + 3) * (H_pdc + H_sfc)
    return gr + 2.0 * H_sfc * nQ + 4.0 * (H_pdc + nQ)


def kernel_roofline(name, avg_s, n, props_per_launch):
    """Algorithmic work of one launch of the named engine kernel."""
    if name == "k_delta":
        flops = OPS_PER_PAIR * 2.0 * n * props_per_launch      # old + new position vs all atoms
    elif name == "k_eval":
        flops = eval_flops_per_proposal() * props_per_launch
    elif name == "k_commit":
        flops = OPS_PER_PAIR * 2.0 * n
    else:
        return None
    tf = flops / avg_s / 1e12
    return {"bound": "mfma", "compute": "fp32 VALU (no matrix instructions; MI355X fp32 vector and dense "
            "matrix peaks are both 157.3 TFLOP/s)", "achieved": round(tf, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tf / PEAK_FP32_TFLOPS, 4), "traffic": None,
            "kernel": name, "avg_launch_us": round(avg_s * 1e6, 2), "flops_per_launch": flops}


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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the pair path runs on a single MI355X; "
              "a sharded multi-GPU run is not implemented.", file=sys.stderr)
        sys.exit(2)

    import torch
    from fullrmc_amd import _lib
    from fullrmc_amd._device import require_gpu
    require_gpu()
    L = _lib.lib()

    start, elem, pdc, sfc = build_workload()
    n = len(elem)
    print("bench: workload built (%d atoms)" % n, file=sys.stderr, flush=True)
    E = make_engine(start, pdc, sfc, seed=1)
    for _ in range(args.warmup):
        E.run_device(args.moves, k=args.k, pull=False)
    h = E.device_handle()
    acc0 = E.accepted
    _lib.check(L.frmc_engine_profile(h, 1))
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
    dom = max(range(nk), key=lambda i: ms[i])
    name = _lib.KERNEL_NAMES[dom]
    avg_s = ms[dom] * 1e-3 / max(1, launches[dom])
    roof = kernel_roofline(name, avg_s, n, evaluated / max(1, windows))

    result = {"metric": "bit-exact-trajectory RMC moves/s + full-histogram Gpairs/s, GPU vs host CPU",
              "value": round(value, 1), "unit": "moves/s", "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3),
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
              "data": "synthetic",
              "config": {"workload": "B2 alloy 13500 atoms, PDC r[0,20]@0.02 + SFC Q[0.5,25]x500, "
                                     "single-atom translation 0.1 A, k=%d speculative windows" % args.k,
                         "atoms": n, "moves_per_step": args.moves, "k": args.k},
              "roofline": roof,
              "accepted_fraction": round(accepted / moves, 4),
              "windows": windows, "proposals_evaluated": evaluated, "kernels": kernels}

    if not args.no_extras:
        # full-histogram recompute (one O(N^2) pass on the PDC grid)
        gps, fms = gpu_full_histogram(start, elem)
        tf = gps * OPS_PER_PAIR / 1e12
        result["full_histogram"] = {"gpairs_per_s": round(gps / 1e9, 3), "kernel_ms": round(fms, 3),
                                    "pairs": n * (n - 1) // 2,
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
        result["plain_one_move_per_window_moves_per_s"] = round(1000 / (time.perf_counter() - t1), 1)
        # the reference's loop structure calling the drop-in extensions per move
        E2 = make_engine(start, pdc, sfc, seed=1)
        E2.run(2)
        t2 = time.perf_counter()
        E2.run(30)
        result["dropin_per_move_loop_moves_per_s"] = round(30 / (time.perf_counter() - t2), 1)
        print("bench: extras done", file=sys.stderr, flush=True)

    if not args.no_cpu:
        cps, cn, cel, cgp = cpu_baseline(start, pdc, sfc, args.cpu_seconds)
        result["cpu_baseline"] = {"value": round(cps, 2), "unit": "moves/s", "cores": 1, "kind": "port",
                                  "sample": "%d moves of the same workload (oracle: C restatement of the reference "
                                            "loops + NumPy chain, 1 thread) in %.1f s" % (cn, cel),
                                  "full_histogram_gpairs_per_s": round(cgp / 1e9, 4)}
        result["speedup_vs_cpu"] = round(value / cps, 1)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
