#!/usr/bin/env python3
"""Benchmark: leaf nodes/sec of batched and served perft on one MI355X.

Headline ("value"): the C2 batch -- 256 seeded random-playout positions
(8..40 plies, mate/stalemate discarded), perft depth 5 each, TT off -- as one
launch per level queued on one stream, every level sized and placed on the
device (pf_perft_batch_mode(..., AUTO) after the autotuner, which picks
this form or the cooperative launch; the one used is in config.mode).
Positions and counts stay in
device memory: a step is the on-device root conversion, the level launches
and the device-side count write-back.  The same steps from host buffers
(PCIe-inclusive) are reported as "host_buffers", and the same batch in one
persistent cooperative launch (grid barriers between levels) and through the
host-driven level loop under "backends".

Also reported on the same box, outside the timed region:
  roofline      integer-VALU roofline of the step's kernels: algorithmic int32
                lane-ops per step (C2 op model x leaves) / the step's kernel
                time (HIP events on the engine's stream around the level
                launches) against MI355X's published vector rate,
                157.3 TFLOP/s FP32 = 78.6 T lane-op/s
  plain         the same 256 positions through the plain per-position path
  backends      batched nodes/sec per slider backend, and the level modes
  serving       C4: the whole configured stream (Poisson arrivals, 50 req/s
                for 120 s, about 6,000 requests, depths {3:.30,4:.30,5:.25,
                6:.12,7:.03}, 32 GiB shared TT) replayed in real time: p50/p99
                latency overall and by depth, nodes/sec, table hits; the same
                requests all at once (serving capacity) and one after another
                through plain pf_perft; every count against the oracle's
  c3            one 20-30-ply middlegame position, perft 8, 64 GiB TT with
                in-level duplicate elimination, against the oracle's perft 8
                (and perft 7, and perft 8 with the TT off); roofline over two
                resources: table traffic at the live random-access and copy
                rates, and the counting work at the vector peak; frac is
                the larger time over the run's (tt_roofline)
  c5            the 4096 high-mobility positions (--c5-positions), perft 6,
                128 GiB TT, every count against the oracle's
  tt_tuning     the table path's launch parameters, tuned on 512 C5 roots
  c1            the initial position, perft 5, through the plain path on the
                GPU (median of 20 runs; gpu_stream: the same as a one-root
                stream-ordered batch); on the CPU below
  cpu_*         the oracle (CPU restatement of the reference perft) on the
                host's CPUs (at most 16 threads, the share one GPU's job gets;
                nproc and the model recorded): C2 in full (fancy magics; and
                one thread), C1 (Kogge-Stone, 50 runs, one thread and split by
                root move), the first two C3 root moves at perft 7, C4's first
                seconds of requests offline, and 64 C5 roots at perft 6
"""
import argparse
import json
import os
import sys
import tempfile
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK_TOPS = 157.3 / 2  # published FP32 vector TFLOP/s / 2 flops per lane-op


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_goldens():
    """Oracle goldens: tests/golden/workloads.json (C2 perft 1-5, C3 perft
    1-7, ...) and, when present, tests/golden/deep.json (C3 perft 8 divided by
    root move, perft 6 of all 4096 C5 positions, every C4 request), the
    latter under the keys c3_deep, c5_deep and c4_deep."""
    gdir = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(gdir, "workloads.json")) as f:
        g = json.load(f)
    deep = os.path.join(gdir, "deep.json")
    if os.path.exists(deep):
        with open(deep) as f:
            d = json.load(f)
        if "perft8" in d.get("c3", {}):
            g["c3_deep"] = d["c3"]
        if "total" in d.get("c5", {}):
            g["c5_deep"] = d["c5"]
        if "requests" in d.get("c4", {}):
            g["c4_deep"] = d["c4"]
    return g


def host_cpus():
    """CPUs this process may run on, the threads the CPU legs use (the
    machine grants one GPU's job a share of at most 16 CPUs), and the model."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "allowed": allowed, "threads": max(1, min(allowed, 16)), "model": model}


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    return Oracle()


def _rows(positions):
    import numpy as np
    from perft_gpu_amd import POS_DTYPE
    return np.array([p.as_tuple() for p in positions], dtype=POS_DTYPE)


def cpu_baseline(positions, depth, gpu_counts, threads, budget_s=20.0):
    """Oracle (plain C restatement of the reference CPU perft, fancy magics)
    on the host cores: test infrastructure used here only as the CPU
    baseline, never as the measured path."""
    orc = _oracle()
    orc.set_backend(2)
    t = time.perf_counter()
    probe = orc.perft_batch(positions[:4], depth - 1, nthreads=1)
    dt = max(time.perf_counter() - t, 1e-6)
    est_leaf_rate = max(float(probe.sum()) / dt, 1e6)
    per_pos = float(gpu_counts.mean())
    n = int(max(threads, min(len(positions), budget_s * est_leaf_rate * threads / max(per_pos, 1.0))))
    n = min(n, len(positions))
    # repeat the sample until about half the budget has elapsed, so that a
    # fast host still yields a timing of several seconds
    t = time.perf_counter()
    reps, nodes, ok = 0, 0, True
    while True:
        got = orc.perft_batch(positions[:n], depth, nthreads=threads)
        reps += 1
        nodes += int(got.sum())
        ok = ok and bool((got == gpu_counts[:n]).all())
        if time.perf_counter() - t >= budget_s / 2 or reps >= 64:
            break
    el = time.perf_counter() - t
    return {"value": nodes / el, "unit": "nodes/s", "cores": threads, "kind": "port",
            "sample": "first %d of the %d C2 positions, perft(%d) each, %d passes, oracle fancy magics, "
                      "%d threads, %.2f s" % (n, len(positions), depth, reps, threads, el),
            "counts_match_gpu": ok,
            "single_thread": {"value": est_leaf_rate, "unit": "nodes/s", "cores": 1,
                              "sample": "first 4 C2 positions, perft(%d) each, one pass" % (depth - 1)}}


def _pcts(times):
    v = sorted(times)
    return {"median_s": v[len(v) // 2], "p10_s": v[int(0.1 * (len(v) - 1))], "p90_s": v[int(0.9 * (len(v) - 1))]}


def cpu_c1(fen, depth, want, threads, reps=50):
    """C1 on the CPU as the configuration names it: the oracle's Kogge-Stone
    perft of one position, `reps` runs each on one thread and on `threads`
    threads (split by root move), median / p10 / p90."""
    orc = _oracle()
    old = orc.set_backend(0)  # Kogge-Stone, for the calling thread (workers inherit it)
    try:
        root = orc.pos(fen)
        kids = _rows(orc.children(root))
        one, many, ok = [], [], True
        for _ in range(reps):
            t = time.perf_counter()
            got = orc.perft(root, depth)
            one.append(time.perf_counter() - t)
            ok = ok and got == want
            t = time.perf_counter()
            got = int(orc.perft_batch(kids, depth - 1, nthreads=threads).sum())
            many.append(time.perf_counter() - t)
            ok = ok and got == want
    finally:
        orc.set_backend(old)
    p1, pm = _pcts(one), _pcts(many)
    return {"value": want / p1["median_s"], "unit": "nodes/s", "runs": reps, "cores": 1, "kind": "port",
            "backend": "ks", "counts_match": ok, **p1,
            "all_cores": {"value": want / pm["median_s"], "unit": "nodes/s", "cores": threads, "runs": reps,
                          "split": "by root move (%d subtrees of perft %d)" % (len(kids), depth - 1), **pm}}


def cpu_c3_subset(c3_pos, golden, threads, subtrees=2):
    """C3 on the CPU: the first `subtrees` root moves' perft 7 (of the 36),
    each split into its perft-6 children over the threads; against the
    golden divide of perft 8."""
    orc = _oracle()
    orc.set_backend(2)
    root = orc.from48(c3_pos)
    moves = orc.moves(root)[:subtrees]
    tasks = []
    for m in moves:
        tasks.extend(orc.children(orc.make(root, m)))
    t = time.perf_counter()
    got = orc.perft_batch(_rows(tasks), 6, nthreads=threads)
    el = time.perf_counter() - t
    nodes = int(got.sum())
    ok = None
    if golden:
        ok = nodes == sum(golden["divide_d8"]["%d" % m] for m in moves)
    return {"value": nodes / el, "unit": "nodes/s", "cores": threads, "kind": "port", "seconds": el,
            "nodes": nodes, "counts_match_golden": ok,
            "sample": "perft 7 of the first %d of the C3 position's %d root moves, split into %d perft-6 subtrees, "
                      "oracle fancy magics" % (len(moves), len(orc.moves(root)), len(tasks))}


def cpu_c4_offline(W, golden, threads, seconds=4.0):
    """C4 offline on the CPU: the requests that arrive in the stream's first
    `seconds`, no pacing, every request split by root move; nodes/s only."""
    orc = _oracle()
    orc.set_backend(2)
    reqs = [r for r in W.serving_requests(W.C4) if r[0] < seconds]
    by_depth, owner = {}, {}
    for i, (_, d, sd) in enumerate(reqs):
        p = orc.from48(W.request_position(sd))
        if d <= 1:
            by_depth.setdefault(0, []).append((i, p))
            continue
        for c in orc.children(p):
            by_depth.setdefault(d - 1, []).append((i, c))
    per = [0] * len(reqs)
    t = time.perf_counter()
    for d, items in sorted(by_depth.items(), reverse=True):
        got = orc.perft_batch(_rows([c for _, c in items]), d, nthreads=threads)
        for (i, _), n in zip(items, got):
            per[i] += int(n)
    el = time.perf_counter() - t
    ok = None
    if golden:
        ok = per == [r[3] for r in golden["requests"][:len(reqs)]]
    return {"value": sum(per) / el, "unit": "nodes/s", "cores": threads, "kind": "port", "seconds": el,
            "nodes": sum(per), "requests": len(reqs), "counts_match_golden": ok,
            "sample": "the %d C4 requests arriving in the first %.0f s, processed offline, split by root move, "
                      "oracle fancy magics" % (len(reqs), seconds)}


def cpu_c5_subset(c5_positions, golden, threads, n=64):
    """C5 on the CPU: perft 6 of the first n of the 4096 positions."""
    orc = _oracle()
    orc.set_backend(2)
    t = time.perf_counter()
    got = orc.perft_batch(c5_positions[:n], 6, nthreads=threads)
    el = time.perf_counter() - t
    ok = None
    if golden:
        ok = [int(x) for x in got] == golden["perft6"][:n]
    return {"value": float(got.sum()) / el, "unit": "nodes/s", "cores": threads, "kind": "port", "seconds": el,
            "nodes": int(got.sum()), "counts_match_golden": ok,
            "sample": "perft 6 of the first %d C5 positions, oracle fancy magics" % n}


def run_serving(pg, W, seconds, backend, paced=True, setup=None):
    """Replay the C4 stream's first `seconds` of arrivals; paced=False submits
    everything at once (capacity); setup(engine) runs before serving starts."""
    cfg = W.C4
    reqs = W.serving_requests(cfg)
    reqs = [r for r in reqs if r[0] < seconds]
    poss = [W.request_position(s) for _, _, s in reqs]
    eng = pg.Engine(backend=backend, arena_bytes=96 << 30, tt_bytes=cfg["tt_bytes"], zobrist_seed=cfg["seed"])
    results = {}
    done = threading.Event()
    try:
        if setup:
            setup(eng)
        eng.serve_start(max_jobs=256)

        def poller():
            while not done.is_set() or len(results) < len(reqs):
                for r in eng.poll():
                    results[r["tag"]] = r
                if len(results) >= len(reqs):
                    break
                time.sleep(0.0002)

        th = threading.Thread(target=poller, daemon=True)
        th.start()
        try:
            t0 = time.perf_counter()
            c0 = eng.now()
            next_log = 10.0
            for i, ((t, d, _), p) in enumerate(zip(reqs, poss)):
                wait = t - (time.perf_counter() - t0)
                if paced and wait > 0:
                    time.sleep(wait)
                eng.submit(p, d, i)
                if paced and t >= next_log:
                    log("serving: %.0f s of arrivals, %d submitted, %d done" % (t, i + 1, len(results)))
                    next_log += 10.0
        finally:
            done.set()
            th.join(timeout=600)
        st = eng.stats()
        eng.serve_stop()
    finally:
        done.set()
        eng.close()
    lat = sorted(r["done_s"] - r["submit_s"] for r in results.values())
    nodes = sum(r["nodes"] for r in results.values())
    span = max(r["done_s"] for r in results.values()) - c0
    bad = sum(1 for r in results.values() if r["status"] != 0)
    by_depth = {}
    for r in results.values():
        by_depth.setdefault(r["depth"], []).append(r["done_s"] - r["submit_s"])

    def pct(v, q):
        return v[min(len(v) - 1, int(q * len(v)))] * 1e3 if v else None

    return {"paced": paced, "requests": len(reqs), "completed": len(results), "errors": bad,
            "p50_ms": pct(lat, 0.50), "p99_ms": pct(lat, 0.99), "max_ms": lat[-1] * 1e3 if lat else None,
            "nodes": nodes, "nodes_per_s": nodes / span if span > 0 else None, "span_s": span,
            "arrival_rate_per_s": cfg["rate"], "seconds": seconds, "steps": st["levels_run"],
            "step_busy_s": st["gpu_ms"] * 1e-3,
            "dup_nodes": st["dup_nodes"], "tt_probes": st["tt_probes"], "tt_hits": st["tt_hits"],
            "requests_by_depth": {str(d): len(v) for d, v in sorted(by_depth.items())},
            "p50_ms_by_depth": {str(d): pct(sorted(v), 0.5) for d, v in sorted(by_depth.items())},
            "p99_ms_by_depth": {str(d): pct(sorted(v), 0.99) for d, v in sorted(by_depth.items())},
            "_counts": [results[i]["nodes"] if i in results else None for i in range(len(reqs))]}


def run_plain_requests(pg, W, seconds, backend):
    """The same request stream through plain per-position perft, one request
    after another (same shared-TT configuration)."""
    cfg = W.C4
    reqs = [r for r in W.serving_requests(cfg) if r[0] < seconds]
    eng = pg.Engine(backend=backend, arena_bytes=96 << 30, tt_bytes=cfg["tt_bytes"], zobrist_seed=cfg["seed"])
    try:
        counts = []
        by_depth = {}
        t = time.perf_counter()
        for _, d, s in reqs:
            t1 = time.perf_counter()
            counts.append(eng.perft(pg.Pos48(*[int(x) for x in W.request_position(s)]), d))
            by_depth[str(d)] = by_depth.get(str(d), 0.0) + time.perf_counter() - t1
        el = time.perf_counter() - t
        nodes = sum(counts)
        st = eng.stats()
        return {"requests": len(reqs), "nodes": nodes, "seconds": el, "nodes_per_s": nodes / el,
                "seconds_by_depth": by_depth, "dup_nodes": st["dup_nodes"], "expand_gpu_s": st["gpu_ms"] * 1e-3,
                "_counts": counts}
    finally:
        eng.close()


def measure_ceilings(pg):
    """Random 16-byte read-modify-writes per second over a 16 GiB table, and
    streaming copy bytes per second (read + write), on this device."""
    eng = pg.Engine(arena_bytes=64 << 20)
    try:
        return {"random16_per_s": eng.probe(2, 16 << 30, 0), "copy_bytes_per_s": eng.probe(1, 1 << 30, 0)}
    finally:
        eng.close()


FRONTIER_NODE_BYTES = 80  # board 48 + root word 4 + two keys 16 + parent 4 + accumulator 8


def issue_view(leaves, kernel_ms, backend):
    """VALU issue rate of a C2 step against the instruction mix's limit (Kogge-Stone only: the per-leaf
    instruction count was measured for its kernels)."""
    from perft_gpu_amd import workloads as W
    if backend != "ks" or kernel_ms <= 0:
        return None
    rate = W.C2_VALU_LANE_INSTR_PER_LEAF * leaves / 64 / (kernel_ms * 1e-3 * 1024 * 2.4e9)
    return {"lane_instr_per_leaf": W.C2_VALU_LANE_INSTR_PER_LEAF, "wave_instr_per_simd_cycle": rate,
            "mix_limit": W.C2_LOOP_ISSUE_LIMIT, "frac": rate / W.C2_LOOP_ISSUE_LIMIT,
            "source": "profiles/k_stream_level_ks_c2_d5_counters.md"}


def tt_roofline(st, seconds, ceil):
    """Roofline of a TT run over its two resources: the table and duplicate
    traffic (HBM random access) and the move counting (integer VALU).
      memory: every probe, store and duplicate claim, and the duplicates'
              two reads in the upsweep, is one random access of at most
              32 bytes, priced at the measured random 16-byte rate; every
              frontier node is written once and read once at streaming rate;
      valu:   every made child that is counted (not a table hit, not a
              duplicate) costs the C2 op model's lane-ops per made node
              (make + count, and generate where it is expanded further),
              priced at the published vector peak.
    The write levels do both at once, so the bound is the larger of the two
    times: frac = max(t_memory, t_valu) / t_run.  serial_frac prices the
    two as if nothing overlapped, (t_memory + t_valu) / t_run; it can pass 1
    because the op model counts two-input operations and the kernels fuse
    them."""
    from perft_gpu_amd.workloads import C2_OPS_PER_MADE_NODE as W_OPS_PER_MADE_NODE
    dups, front = st["dup_nodes"], st["frontier_nodes"]
    rnd = st["tt_probes"] + st["tt_stores"] + front + 3 * dups
    stream = 2 * FRONTIER_NODE_BYTES * front
    t_mem = rnd / ceil["random16_per_s"] + stream / ceil["copy_bytes_per_s"]
    counted = max(0, st["made_nodes"] - dups - st["tt_hits"])
    t_valu = counted * W_OPS_PER_MADE_NODE / (VALU_PEAK_TOPS * 1e12)
    return {"bound": "valu" if t_valu >= t_mem else "hbm-random", "random_accesses": rnd, "stream_bytes": stream,
            "t_memory_s": t_mem, "counted_nodes": counted, "ops_per_counted_node": W_OPS_PER_MADE_NODE,
            "t_valu_s": t_valu, "t_min_s": max(t_mem, t_valu), "random16_per_s": ceil["random16_per_s"],
            "copy_bytes_per_s": ceil["copy_bytes_per_s"], "memory_frac": t_mem / seconds,
            "frac": max(t_mem, t_valu) / seconds, "serial_frac": (t_mem + t_valu) / seconds}


def run_tt_workload(pg, W, cfg, positions, depth, backend, arena, check=None, ceil=None, tt=True, tune_file=None,
                    setup=None):
    eng = pg.Engine(backend=backend, arena_bytes=arena, tt_bytes=cfg["tt_bytes"] if tt else 0,
                    zobrist_seed=cfg["seed"], tune_file=tune_file)
    try:
        if setup:
            setup(eng)
        t = time.perf_counter()
        got = eng.perft_batch(positions, depth)
        el = time.perf_counter() - t
        st = eng.stats()
        out = {"depth": depth, "positions": len(positions), "nodes": int(got.sum()), "seconds": el,
               "nodes_per_s": float(got.sum()) / el, "tt_gib": (cfg["tt_bytes"] >> 30) if tt else 0,
               "tt_probes": st["tt_probes"], "tt_hits": st["tt_hits"], "tt_stores": st["tt_stores"],
               "dup_nodes": st["dup_nodes"], "frontier_nodes": st["frontier_nodes"], "made_nodes": st["made_nodes"],
               "chunks": st["chunks"], "graph_passes": st["graph_passes"],
               "peak_arena_gib": st["peak_arena_bytes"] / 2 ** 30}
        if tt and ceil:
            out["roofline"] = tt_roofline(st, el, ceil)
        if check is not None:
            out["counts_match"] = bool(check(got))
        return out
    finally:
        eng.close()


def _children(pg, eng, positions, ep_every_double_push):
    """Every legal child of every position (as pf_pos48 tuples), with the en
    passant file set the way the run that stored them sets it."""
    kids = []
    for r in positions:
        pos = pg.Pos48(*[int(x) for x in r])
        for uci, _ in eng.divide(pos, 1):
            c = pg.apply_uci(pos, uci)
            if ep_every_double_push:
                frm = (ord(uci[0]) - 97) + 8 * (ord(uci[1]) - 49)
                to = (ord(uci[2]) - 97) + 8 * (ord(uci[3]) - 49)
                pawns = pos.pawns & 0x00FFFFFFFFFFFF00  # ranks 2-7; the other bits hold the state
                if (pawns >> frm) & 1 and abs(to - frm) == 16:
                    c.pawns = (c.pawns & ~0xF0) | (((frm & 7) + 1) << 4)
            kids.append(c.as_tuple())
    return kids


def table_semantics(pg, W, backend, c3, want8, tune_file):
    """The table's semantics measured (the reference's switches.h:43-48 and
    launcher.h:62-126):
      c3_d8    C3 perft 8 with exact en-passant flagging and with the file set
               after every double push: in-level duplicates found, nodes
               written and made, time;
      requery  256 C5 positions at perft 5 on a 4 MiB table (2^17 buckets,
               far fewer than the entries stored), then all their children
               at perft 4 on the same table: the second batch's probes and
               hits, with one depth-keyed table, with four per-depth regions,
               and with always-set flagging."""
    out = {"c3_d8": {}, "requery": {}}
    for name, exact in (("exact_ep", True), ("ep_every_double_push", False)):
        r = run_tt_workload(pg, W, W.C3, c3, 8, backend, 128 << 30, tune_file=tune_file,
                            check=(lambda got: int(got[0]) == want8) if want8 else None,
                            setup=lambda e, x=exact: e.set_exact_ep(x))
        out["c3_d8"][name] = {k: r[k] for k in ("seconds", "dup_nodes", "frontier_nodes", "made_nodes",
                                                "tt_probes", "tt_hits", "counts_match") if k in r}
    roots = W.positions(W.C5)[:256]
    for name, exact, regions in (("one_table", True, 0), ("per_depth_regions_4", True, 2),
                                 ("ep_every_double_push", False, 0)):
        eng = pg.Engine(backend=backend, arena_bytes=8 << 30, tt_bytes=4 << 20, zobrist_seed=W.C5["seed"])
        try:
            eng.set_exact_ep(exact)
            eng.set_tt_regions(regions)
            first = eng.perft_batch(roots, 5)
            kids = _children(pg, eng, roots, not exact)
            eng.reset_stats()
            second = eng.perft_batch(kids, 4)
            st = eng.stats()
            out["requery"][name] = {"first_nodes": int(first.sum()), "children": len(kids),
                                    "second_nodes": int(second.sum()), "probes": st["tt_probes"],
                                    "hits": st["tt_hits"], "hit_rate": st["tt_hits"] / max(1, st["tt_probes"]),
                                    "made_nodes": st["made_nodes"]}
        finally:
            eng.close()
    same = len(set((v["first_nodes"], v["second_nodes"]) for v in out["requery"].values())) == 1
    out["requery_counts_agree"] = same
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--backend", default="ks")
    ap.add_argument("--serve-seconds", type=float, default=120.0,
                    help="seconds of the C4 arrival stream to serve (120 = the whole configured stream)")
    ap.add_argument("--c5-positions", type=int, default=4096)
    ap.add_argument("--no-extras", action="store_true", help="skip every comparison run")
    ap.add_argument("--sections", default="batch,serving,tt",
                    help="comparison runs to do: batch (host buffers, plain, backends), serving, tt (C3, C5)")
    ap.add_argument("--serve-modes", default="paced,capacity,plain",
                    help="serving runs: paced (real-time arrivals), capacity (all at once), plain (per request)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the built-in launch shape")
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
    gold = load_goldens()
    eng = pg.Engine(backend=args.backend, arena_bytes=16 << 30)
    log("bench: %d positions depth %d backend %s" % (len(pos), depth, args.backend))
    tuning = None
    import shutil
    import tempfile
    tune_dir = tempfile.mkdtemp()
    tf = os.path.join(tune_dir, "tune.txt")
    if not args.no_tune:
        # the autotuner times every launch shape of the table-free passes on
        # this workload and keeps the fastest
        t = time.perf_counter()
        eng.tune(pos, depth, tf, what=pg.TUNE_SINGLE_PASS)
        tuning = {"seconds": time.perf_counter() - t,
                  "choice": [l.strip() for l in open(tf) if l.startswith("backend " + args.backend + " ")]}
        log("tuning: %s" % tuning)
    # the single pass PF_MODE_AUTO runs for this backend (stream until tuned)
    single = "stream"
    if tuning and tuning["choice"] and "single_pass persistent" in tuning["choice"][0]:
        single = "persistent"
    # positions and counts resident in HBM: a step is the on-device conversion
    # of the roots, the level launches and the count write-back
    dpos = pg.DeviceArray.from_host(pos)
    dout = pg.DeviceArray(len(pos), np.uint64)
    for i in range(args.warmup):
        eng.perft_batch_ptr(dpos.ptr, len(pos), depth, dout.ptr, pg.MODE_AUTO)
        log("warmup %d done" % i)
    eng.sync()
    eng.reset_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.perft_batch_ptr(dpos.ptr, len(pos), depth, dout.ptr, pg.MODE_AUTO)
    eng.sync()
    el = time.perf_counter() - t0
    st = eng.stats()
    counts = dout.download()
    dpos.free()
    dout.free()
    leaves = int(counts.sum())
    value = leaves * args.steps / el
    # per step: the root conversion, then the root count and depth - 1 level
    # launches; gpu_ms holds the span from the root count to the last level
    kernel_ms = st["gpu_ms"] / args.steps
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
                   "backend": args.backend, "mode": "auto: " + single, "buffers": "device",
                   "leaves_per_step": leaves,
                   # every position's count against the oracle's golden perft 5
                   "counts_match_golden": counts.tolist() == gold["c2"]["perft"][str(depth)],
                   "launch_form": ("one launch per level queued on one stream, each level sized and placed on "
                                   "the device (the same batch in one persistent cooperative launch is "
                                   "backends.persistent_mode_%s)" % args.backend) if single == "stream" else
                                  "one persistent cooperative launch",
                   "tuning": tuning},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "Tlane-op/s",
                     "frac": achieved / VALU_PEAK_TOPS,
                     # HBM bytes per step: FETCH_SIZE x2.00 + WRITE_SIZE x1.00 over the step's kernels
                     # (rocprofv3 --pmc, calibrated on a copy of known size;
                     # profiles/k_stream_level_ks_c2_d5_counters.md)
                     "traffic": 9.22e8,
                     "traffic_source": "profiles/k_stream_level_ks_c2_d5_counters.md (rocprofv3 --pmc "
                                       "FETCH_SIZE, WRITE_SIZE over one step, calibrated on a copy)",
                     "kernel": ("k_stream_roots + %d x k_stream_level<%s>" % (depth - 1, args.backend)
                                if single == "stream" else "k_persistent<%s>" % args.backend),
                     "kernel_ms": kernel_ms,
                     "ops_per_leaf": W.C2_OPS_PER_LEAF,
                     # the hardware view: VALU wave-instructions issued per SIMD-cycle over the
                     # step (counted instructions per leaf, this run's kernel time, 1,024 SIMDs at
                     # 2.4 GHz) against what the leaf loop's instruction mix can issue
                     "issue": issue_view(leaves, kernel_ms, args.backend)},
    }

    sections = set() if args.no_extras else set(s for s in args.sections.split(",") if s)
    if "batch" in sections:
        # the same steps from host buffers: host-side conversion, 256 x 48-B
        # upload and the 2 KiB count read-back over PCIe on every step
        t = time.perf_counter()
        for _ in range(args.steps):
            c = eng.perft_batch(pos, depth, mode=pg.MODE_AUTO)
        th = time.perf_counter() - t
        line["host_buffers"] = {"value": leaves * args.steps / th, "unit": "nodes/s",
                                "ms_per_step": th * 1e3 / args.steps, "counts_match": bool((c == counts).all())}
        log("host buffers: %.3e nodes/s" % line["host_buffers"]["value"])
        eng.reset_stats()
        t = time.perf_counter()
        plain = np.array([eng.perft(pg.Pos48(*[int(x) for x in r]), depth) for r in pos], dtype=np.uint64)
        tp = time.perf_counter() - t
        line["plain"] = {"value": leaves / tp, "unit": "nodes/s", "ms_per_step": tp * 1e3,
                         "counts_match_batched": bool((plain == counts).all())}
        line["batched_over_plain"] = value / (leaves / tp)
        log("plain: %.3e nodes/s  ratio %.2f" % (leaves / tp, line["batched_over_plain"]))
        # C1: one position (the initial position) at depth 5 through pf_perft
        eng.set_backend(W.C1["backend"])
        c1pos = pg.fen_to_pos(W.C1["fen"])
        c1t = []
        for _ in range(21):
            t = time.perf_counter()
            c1n = eng.perft(c1pos, W.C1["depth"])
            c1t.append(time.perf_counter() - t)
        c1med = sorted(c1t[1:])[len(c1t[1:]) // 2]
        line["c1"] = {"gpu": {"value": c1n / c1med, "unit": "nodes/s", "median_s": c1med, "runs": 20,
                              "nodes": c1n, "counts_match": c1n == 4865609}}
        # the same position as a one-root batch in the stream-ordered mode:
        # no host round trip between levels
        c1s = []
        for _ in range(21):
            t = time.perf_counter()
            c1b = int(eng.perft_batch([c1pos], W.C1["depth"], mode=pg.MODE_STREAM)[0])
            c1s.append(time.perf_counter() - t)
        c1smed = sorted(c1s[1:])[len(c1s[1:]) // 2]
        line["c1"]["gpu_stream"] = {"value": c1b / c1smed, "unit": "nodes/s", "median_s": c1smed, "runs": 20,
                                    "counts_match": c1b == 4865609}
        eng.set_backend(args.backend)
        log("c1: %s" % line["c1"])
        be = {}
        for b in ("ks", "plain", "fancy", "byte"):
            eng.set_backend(b)
            eng.perft_batch(pos, depth, mode=pg.MODE_AUTO)
            t = time.perf_counter()
            reps = 3
            for _ in range(reps):
                c = eng.perft_batch(pos, depth, mode=pg.MODE_AUTO)
            be[b] = {"value": leaves * reps / (time.perf_counter() - t), "counts_match": bool((c == counts).all())}
            log("backend %s: %.3e nodes/s" % (b, be[b]["value"]))
        eng.set_backend(args.backend)
        # the byte backend's other form: its tables staged in each workgroup's LDS
        with tempfile.TemporaryDirectory() as td:
            lds_tf = os.path.join(td, "byte_lds.txt")
            with open(lds_tf, "w") as f:
                f.write("perft_gpu_amd tuning 1\nbackend byte stream_leaf 6 stream_write 6\n")
            e2 = pg.Engine(backend="byte", arena_bytes=16 << 30, tune_file=lds_tf)
            try:
                e2.perft_batch(pos, depth, mode=pg.MODE_STREAM)
                t = time.perf_counter()
                for _ in range(3):
                    c = e2.perft_batch(pos, depth, mode=pg.MODE_STREAM)
                be["byte_tables_in_lds"] = {"value": leaves * 3 / (time.perf_counter() - t),
                                            "counts_match": bool((c == counts).all())}
            finally:
                e2.close()
        log("backend byte, tables in LDS: %.3e nodes/s" % be["byte_tables_in_lds"]["value"])
        for name, mode in (("stream_mode_", pg.MODE_STREAM), ("persistent_mode_", pg.MODE_PERSISTENT),
                           ("levels_mode_", pg.MODE_LEVELS)):
            eng.perft_batch(pos, depth, mode=mode)
            t = time.perf_counter()
            for _ in range(3):
                c = eng.perft_batch(pos, depth, mode=mode)
            be[name + args.backend] = {"value": leaves * 3 / (time.perf_counter() - t),
                                       "counts_match": bool((c == counts).all())}
        line["backends"] = be
    eng.close()

    if "serving" in sections:
        try:
            modes = set(m for m in args.serve_modes.split(",") if m)
            runs = {}
            if "paced" in modes:
                runs["serving"] = run_serving(pg, W, args.serve_seconds, args.backend)
                log("serving: p50 %.3f ms p99 %.3f ms" % (runs["serving"]["p50_ms"], runs["serving"]["p99_ms"]))
            if "capacity" in modes:
                runs["serving_capacity"] = run_serving(pg, W, args.serve_seconds, args.backend, paced=False)
            if "plain" in modes:
                runs["plain_same_requests"] = run_plain_requests(pg, W, args.serve_seconds, args.backend)
            req_counts = {k: r.pop("_counts") for k, r in runs.items()}
            if "c4_deep" in gold:  # the oracle's count of every request of the stream
                want = [r[3] for r in gold["c4_deep"]["requests"]]
                for k, c in req_counts.items():
                    runs[k]["counts_match_golden"] = c == want[:len(c)]
            if "plain_same_requests" in runs:
                for k in ("serving", "serving_capacity"):
                    if k in runs:
                        runs[k]["counts_match_plain"] = req_counts[k] == req_counts["plain_same_requests"]
                if "serving_capacity" in runs:
                    cap, pl = runs["serving_capacity"], runs["plain_same_requests"]
                    line["served_over_plain"] = cap["nodes_per_s"] / pl["nodes_per_s"]
                    log("serving capacity %.3e nodes/s vs plain %.3e" % (cap["nodes_per_s"], pl["nodes_per_s"]))
            line.update(runs)
        except Exception as e:  # report, do not lose the headline
            line["serving_error"] = repr(e)
    if "tt" in sections:
        try:
            g = gold
            ceil = measure_ceilings(pg)
            log("ceilings: %s" % ceil)
            tt_tune = None
            if not args.no_tune:
                # the table path's workgroups per CU and chunk share, tuned on
                # 512 of the C5 positions at perft 6 with C5's table and arena
                c5s = W.positions(W.C5)[:512]
                teng = pg.Engine(backend=args.backend, arena_bytes=96 << 30, tt_bytes=W.C5["tt_bytes"],
                                 zobrist_seed=W.C5["seed"], tune_file=tf if os.path.exists(tf) else None)
                try:
                    t = time.perf_counter()
                    teng.tune(c5s, 6, tf, what=pg.TUNE_TABLE)
                    tt_tune = {"seconds": time.perf_counter() - t, "sample": "first 512 C5 positions, perft 6",
                               "choice": [l.strip() for l in open(tf)
                                          if l.startswith("backend " + args.backend + " ")]}
                finally:
                    teng.close()
                log("table tuning: %s" % tt_tune)
                line["tt_tuning"] = tt_tune
            ttf = tf if os.path.exists(tf) else None
            c3 = W.positions(W.C3)
            want8 = g["c3_deep"]["perft8"] if "c3_deep" in g else None
            line["c3"] = run_tt_workload(pg, W, W.C3, c3, 8, args.backend, 128 << 30, ceil=ceil,
                                         check=(lambda got: int(got[0]) == want8) if want8 else None,
                                         tune_file=ttf)
            line["c3"]["d7"] = run_tt_workload(pg, W, W.C3, c3, 7, args.backend, 128 << 30, ceil=ceil,
                                               check=lambda got: int(got[0]) == g["c3"]["perft"]["7"][0],
                                               tune_file=ttf)
            line["c3"]["d8_tt_off"] = run_tt_workload(pg, W, W.C3, c3, 8, args.backend, 128 << 30, tt=False)
            line["c3"]["speedup_over_tt_off"] = line["c3"]["d8_tt_off"]["seconds"] / line["c3"]["seconds"]
            line["c3"]["counts_match_tt_off"] = line["c3"]["nodes"] == line["c3"]["d8_tt_off"]["nodes"]
            # the same run with the host-driven levels (no captured passes)
            line["c3"]["d8_host_levels"] = run_tt_workload(
                pg, W, W.C3, c3, 8, args.backend, 128 << 30, tune_file=ttf,
                check=(lambda got: int(got[0]) == want8) if want8 else None,
                setup=lambda e: e.set_level_graphs(False))
            log("c3: %s" % line["c3"])
            if args.c5_positions > 0:
                c5 = W.positions(W.C5)[:args.c5_positions]
                want6 = g["c5_deep"]["perft6"] if "c5_deep" in g else g["c5"]["perft_first8_d6"]
                line["c5"] = run_tt_workload(
                    pg, W, W.C5, c5, 6, args.backend, 96 << 30, ceil=ceil,
                    check=lambda got: [int(x) for x in got[:len(want6)]] == want6[:len(got)], tune_file=ttf)
                line["c5"]["checked_positions"] = min(len(c5), len(want6))
                if len(c5) == 4096:
                    line["c5"]["host_levels"] = run_tt_workload(
                        pg, W, W.C5, c5, 6, args.backend, 96 << 30, tune_file=ttf,
                        check=lambda got: [int(x) for x in got[:len(want6)]] == want6[:len(got)],
                        setup=lambda e: e.set_level_graphs(False))
                    log("c5 host levels: %.3f s" % line["c5"]["host_levels"]["seconds"])
                log("c5: %s" % line["c5"])
            line["f2"] = table_semantics(pg, W, args.backend, c3, want8, ttf)
            log("f2: %s" % line["f2"])
        except Exception as e:
            line["tt_runs_error"] = repr(e)
    if not args.no_cpu:
        hc = host_cpus()
        th = hc["threads"]
        line["cpu_baseline"] = cpu_baseline(pos, depth, counts, th)
        line["cpu_baseline"]["host"] = hc
        log("cpu: %s" % line["cpu_baseline"])
        c1cpu = cpu_c1(W.C1["fen"], W.C1["depth"], 4865609, th)
        line.setdefault("c1", {})["cpu"] = c1cpu
        if "gpu" in line["c1"]:
            line["c1"]["gpu_over_cpu"] = line["c1"]["gpu"]["value"] / c1cpu["value"]
            line["c1"]["gpu_over_cpu_all_cores"] = line["c1"]["gpu"]["value"] / c1cpu["all_cores"]["value"]
        log("c1 cpu: %s" % c1cpu)
        line["cpu_c3"] = cpu_c3_subset(W.positions(W.C3)[0], gold.get("c3_deep"), th)
        log("c3 cpu: %s" % line["cpu_c3"])
        line["cpu_c4"] = cpu_c4_offline(W, gold.get("c4_deep"), th)
        log("c4 cpu: %s" % line["cpu_c4"])
        line["cpu_c5"] = cpu_c5_subset(W.positions(W.C5), gold.get("c5_deep"), th)
        log("c5 cpu: %s" % line["cpu_c5"])
    shutil.rmtree(tune_dir, ignore_errors=True)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
