#!/usr/bin/env python3
"""Benchmark: leaf nodes/sec of batched perft on one MI355X.

Workload (one "step"): the C2 batch -- 256 seeded random-playout positions
(8..40 plies, mate/stalemate discarded), perft depth 5 each, TT off, in one
persistent cooperative launch (pf_perft_batch_mode(..., PF_MODE_PERSISTENT)).
The positions live in HBM only for the duration of a call, so every step
uploads 256 x 48 bytes; that upload is inside the timed region.

Printed JSON line (rank 0 only):
  value         leaf nodes/sec of the batched path over the timed steps
  roofline      integer-VALU roofline of the persistent kernel: algorithmic
                int32 lane-ops per launch (op model x leaves) / launch time,
                against 256 CU x 64 lanes x 2.4 GHz = 39.3 T lane-op/s
  plain         the same positions through the plain per-position path
  backends      batched nodes/sec per slider backend
  serving       p50/p99 request latency + nodes/sec of the serving loop
  cpu_baseline  the oracle (CPU restatement of the reference perft, fancy
                magics) on the host cores, same positions, same depth
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # int32 lane-ops/s, nominal MI355X


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(positions, depth, gpu_counts, budget_s=20.0):
    """Oracle (plain C restatement of the reference CPU perft, fancy magics)
    on the host cores: test infrastructure used here only as the CPU
    baseline, never as the measured path."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    orc = Oracle()
    orc.set_backend(2)
    threads = max(1, min(os.cpu_count() or 1, 16))
    # size the sample to ~budget_s of CPU work from a quick single-thread probe
    t = time.perf_counter()
    probe = orc.perft_batch(positions[:4], depth - 1, nthreads=1)
    dt = max(time.perf_counter() - t, 1e-6)
    est_leaf_rate = max(float(probe.sum()) / dt, 1e6)
    per_pos = float(gpu_counts.mean())
    n = int(max(threads, min(len(positions), budget_s * est_leaf_rate * threads / max(per_pos, 1.0))))
    n = min(n, len(positions))
    t = time.perf_counter()
    got = orc.perft_batch(positions[:n], depth, nthreads=threads)
    el = time.perf_counter() - t
    match = bool((got == gpu_counts[:n]).all())
    return {"value": float(got.sum()) / el, "unit": "nodes/s", "cores": threads, "kind": "port",
            "sample": "first %d of the %d positions, perft(%d) each, oracle fancy magics, %d threads, %.2f s"
                      % (n, len(positions), depth, threads, el),
            "counts_match_gpu": match}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--backend", default="ks")
    ap.add_argument("--no-extras", action="store_true", help="skip plain/backend/serving/cpu comparisons")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU only; --gpus %d is not supported (use --gpus 1)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import perft_gpu_amd as pg
    from perft_gpu_amd import workloads as W

    cfg = W.C2
    depth = cfg["depth"]
    pos = W.positions(cfg)
    eng = pg.Engine(backend=args.backend, arena_bytes=16 << 30)
    log("bench: %d positions depth %d backend %s" % (len(pos), depth, args.backend))
    for i in range(args.warmup):
        counts = eng.perft_batch(pos, depth, mode=pg.MODE_PERSISTENT)
        log("warmup %d done" % i)
    eng.sync()
    eng.reset_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        counts = eng.perft_batch(pos, depth, mode=pg.MODE_PERSISTENT)
    eng.sync()
    el = time.perf_counter() - t0
    st = eng.stats()
    leaves = int(counts.sum())
    value = leaves * args.steps / el
    kernel_ms = st["gpu_ms"] / max(st["kernel_launches"], 1)
    ops = W.C2_OPS_PER_LEAF * leaves
    achieved = ops / (kernel_ms * 1e-3) / 1e12
    log("timed: %.3f s, %.3e nodes/s, kernel %.3f ms" % (el, value, kernel_ms))

    line = {
        "metric": "leaf nodes/sec, batched perft (C2: 256 playout positions x perft 5, TT off), 1 MI355X",
        "value": value,
        "unit": "nodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: seeded random legal playouts from the initial position",
        "config": {"workload": cfg["name"], "positions": len(pos), "depth": depth, "tt": "off",
                   "backend": args.backend, "mode": "persistent", "leaves_per_step": leaves},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "Tlane-op/s",
                     "frac": achieved / VALU_PEAK_TOPS, "traffic": None,
                     "kernel": "k_persistent<%s>" % args.backend, "kernel_ms": kernel_ms,
                     "ops_per_leaf": W.C2_OPS_PER_LEAF},
    }

    if not args.no_extras:
        # plain per-position perft: the same positions one at a time
        eng.reset_stats()
        t = time.perf_counter()
        plain = np.array([eng.perft(pg.Pos48(*[int(x) for x in r]), depth) for r in pos], dtype=np.uint64)
        tp = time.perf_counter() - t
        line["plain"] = {"value": leaves / tp, "unit": "nodes/s", "ms_per_step": tp * 1e3,
                         "counts_match_batched": bool((plain == counts).all())}
        line["batched_over_plain"] = value / (leaves / tp)
        log("plain: %.3e nodes/s  ratio %.2f" % (leaves / tp, line["batched_over_plain"]))
        # backends and the host-driven level mode
        be = {}
        for b in ("ks", "plain", "fancy", "byte"):
            eng.set_backend(b)
            eng.perft_batch(pos, depth, mode=pg.MODE_PERSISTENT)
            t = time.perf_counter()
            reps = 3
            for _ in range(reps):
                c = eng.perft_batch(pos, depth, mode=pg.MODE_PERSISTENT)
            be[b] = {"value": leaves * reps / (time.perf_counter() - t), "counts_match": bool((c == counts).all())}
            log("backend %s: %.3e nodes/s" % (b, be[b]["value"]))
        eng.set_backend(args.backend)
        t = time.perf_counter()
        c = eng.perft_batch(pos, depth, mode=pg.MODE_LEVELS)
        be["levels_mode_" + args.backend] = {"value": leaves / (time.perf_counter() - t),
                                             "counts_match": bool((c == counts).all())}
        line["backends"] = be
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(pos, depth, counts)
        log("cpu: %s" % line["cpu_baseline"])
    eng.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
