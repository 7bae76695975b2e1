"""Deep goldens at the configured sizes, from the oracle (CPU restatement of
the reference perft; test infrastructure), split over the host's cores:

  c3   the C3 position: perft 7 of every legal root move (its divide at depth
       8) and their sum, perft 8 = 1,531,176,662,801 leaves;
  c5   perft 6 of all 4096 C5 positions (about 8.1e12 leaves);
  c4   the perft of every request of the 120 s C4 stream (depths 3..7).

Work is checkpointed into tests/golden/deep.json after every piece, so an
interrupted run resumes where it stopped.  About 40 minutes on 8 cores.

    nice python tests/golden/make_deep_golden.py [--threads N] [--only c3,c5,c4]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle_lib import Oracle  # noqa: E402
from perft_gpu_amd import workloads as W  # noqa: E402

OUT = os.path.join(HERE, "deep.json")


def load():
    if os.path.exists(OUT):
        with open(OUT) as f:
            return json.load(f)
    return {"generated_by": "tests/golden/make_deep_golden.py (oracle, fancy-magic backend)"}


def save(g):
    tmp = OUT + ".tmp"
    with open(tmp, "w") as f:
        json.dump(g, f, separators=(",", ":"))
    os.replace(tmp, OUT)


def rows_of(o, positions):
    return np.array([p.as_tuple() for p in positions],
                    dtype=[("white", "<u8"), ("pawns", "<u8"), ("knights", "<u8"),
                           ("bishop_queens", "<u8"), ("rook_queens", "<u8"), ("kings", "<u8")])


def do_c3(o, g, nthreads):
    c = W.C3
    root = o.playouts(c["seed"], c["n"], c["min_plies"], c["max_plies"], c["min_moves"], c["no_check"])[0]
    rp = o.from48(root)
    moves = o.moves(rp)
    sec = g.setdefault("c3", {"config": c["name"], "divide_d8": {}})
    t0 = time.time()
    for m in moves:
        key = "%d" % m
        if key in sec["divide_d8"]:
            continue
        child = o.make(rp, m)  # generateBoards' flagging: counts do not depend on it
        grand = rows_of(o, o.children(child))
        n = int(o.perft_batch(grand, 6, nthreads=nthreads).sum()) if len(grand) else 0
        sec["divide_d8"][key] = n
        save(g)
        print("c3 move %s: %d (%d/%d, %.0fs)" % (key, n, len(sec["divide_d8"]), len(moves), time.time() - t0),
              flush=True)
    sec["moves"] = ["%d" % m for m in moves]
    sec["perft8"] = sum(sec["divide_d8"].values())
    save(g)
    print("c3 perft 8 = %d" % sec["perft8"], flush=True)


def do_c5(o, g, nthreads, chunk=256):
    c = W.C5
    rows = o.playouts(c["seed"], c["n"], c["min_plies"], c["max_plies"], c["min_moves"], c["no_check"])
    sec = g.setdefault("c5", {"config": c["name"], "n": len(rows), "perft6": []})
    t0 = time.time()
    while len(sec["perft6"]) < len(rows):
        a = len(sec["perft6"])
        got = o.perft_batch(rows[a:a + chunk], 6, nthreads=nthreads)
        sec["perft6"].extend(int(x) for x in got)
        save(g)
        print("c5 %d/%d (%.0fs)" % (len(sec["perft6"]), len(rows), time.time() - t0), flush=True)
    sec["total"] = sum(sec["perft6"])
    save(g)


def do_c4(o, g, nthreads, chunk=512):
    reqs = W.serving_requests(W.C4)
    sec = g.setdefault("c4", {"config": W.C4["name"], "n": len(reqs), "perft": {}})
    todo = [i for i in range(len(reqs)) if "%d" % i not in sec["perft"]]
    # deepest first, so that the threads finish together
    todo.sort(key=lambda i: -reqs[i][1])
    t0 = time.time()
    for a in range(0, len(todo), chunk):
        idx = todo[a:a + chunk]
        by_depth = {}
        for i in idx:
            by_depth.setdefault(reqs[i][1], []).append(i)
        for d, ii in sorted(by_depth.items(), reverse=True):
            rows = np.concatenate([o.playouts(reqs[i][2], 1, W.C4["min_plies"], W.C4["max_plies"], 0, False)
                                   for i in ii])
            got = o.perft_batch(rows, d, nthreads=nthreads)
            for i, n in zip(ii, got):
                sec["perft"]["%d" % i] = int(n)
        save(g)
        print("c4 %d/%d (%.0fs)" % (len(sec["perft"]), len(reqs), time.time() - t0), flush=True)
    sec["requests"] = [[reqs[i][0], reqs[i][1], reqs[i][2], sec["perft"]["%d" % i]] for i in range(len(reqs))]
    del sec["perft"]
    save(g)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--only", default="c3,c5,c4")
    args = ap.parse_args()
    o = Oracle()
    o.set_backend(2)  # fancy magics: the reference default
    g = load()
    parts = args.only.split(",")
    if "c3" in parts and "perft8" not in g.get("c3", {}):
        do_c3(o, g, args.threads)
    if "c4" in parts and "requests" not in g.get("c4", {}):
        do_c4(o, g, args.threads)
    if "c5" in parts and "total" not in g.get("c5", {}):
        do_c5(o, g, args.threads)
    print("done", flush=True)


if __name__ == "__main__":
    main()
