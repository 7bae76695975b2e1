"""Regenerate tests/golden/workloads.json with the oracle (CPU restatement of
the reference perft; test infrastructure).  Positions come from the ORACLE's
own playout generator, so the product's generator is checked against them.

    python tests/golden/make_golden.py          # ~1-3 minutes on 8 cores
"""
import concurrent.futures as cf
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle_lib import Oracle  # noqa: E402
from perft_gpu_amd import workloads as W  # noqa: E402

NTHREADS = min(8, os.cpu_count() or 1)


def hexpos(row):
    return ["%016x" % int(x) for x in row]


def counts(o, rows, depths):
    out = {}
    for d in depths:
        out[str(d)] = [int(x) for x in o.perft_batch(rows, d, nthreads=NTHREADS)]
    return out


def main():
    o = Oracle()
    o.set_backend(2)  # fancy magics: the reference default
    t0 = time.time()
    g = {"generated_by": "tests/golden/make_golden.py (oracle, fancy-magic backend)"}

    c = W.C2
    rows = o.playouts(c["seed"], c["n"], c["min_plies"], c["max_plies"], c["min_moves"], c["no_check"])
    g["c2"] = {"config": c["name"], "positions": [hexpos(r) for r in rows], "perft": counts(o, rows, range(1, 6))}
    print("c2 done %.1fs total leaves d5 %d" % (time.time() - t0, sum(g["c2"]["perft"]["5"])), flush=True)

    # op model over the same workload (Kogge-Stone weights)
    def one(i):
        return o.opcount(o.from48(rows[i]), c["depth"])
    with cf.ThreadPoolExecutor(NTHREADS) as ex:
        res = list(ex.map(one, range(len(rows))))
    ops = sum(r[0] for r in res)
    leaves = sum(r[1] for r in res)
    assert leaves == sum(g["c2"]["perft"]["5"])
    g["c2"]["op_model"] = {"int32_lane_ops": ops, "leaves": leaves, "ops_per_leaf": ops / leaves}
    print("op model %.2f ops/leaf %.1fs" % (ops / leaves, time.time() - t0), flush=True)

    c = W.C3
    rows = o.playouts(c["seed"], c["n"], c["min_plies"], c["max_plies"], c["min_moves"], c["no_check"])
    g["c3"] = {"config": c["name"], "positions": [hexpos(r) for r in rows], "perft": counts(o, rows, range(1, 7))}
    # depth 7 by root move: the children at depth 6, spread over the threads
    import numpy as np
    kids = o.children(o.from48(rows[0]))
    karr = np.array([k.as_tuple() for k in kids], dtype=rows.dtype)
    g["c3"]["perft"]["7"] = [int(o.perft_batch(karr, 6, nthreads=NTHREADS).sum())]
    print("c3 done %.1fs" % (time.time() - t0), flush=True)

    c = W.C5
    rows = o.playouts(c["seed"], c["n"], c["min_plies"], c["max_plies"], c["min_moves"], c["no_check"])
    g["c5"] = {"config": c["name"], "n": len(rows),
               "positions_first64": [hexpos(r) for r in rows[:64]],
               "perft_first64": counts(o, rows[:64], range(1, 6)),
               "perft_first8_d6": [int(x) for x in o.perft_batch(rows[:8], 6, nthreads=NTHREADS)]}
    print("c5 done %.1fs" % (time.time() - t0), flush=True)

    reqs = W.serving_requests(W.C4, limit=300)
    rrows = [o.playouts(s, 1, W.C4["min_plies"], W.C4["max_plies"], 0, False)[0] for _, _, s in reqs]
    import numpy as np
    rarr = np.array(rrows, dtype=rrows[0].dtype)
    rc = []
    for i, (t, d, s) in enumerate(reqs):
        rc.append(int(o.perft_batch(rarr[i:i + 1], d, nthreads=1)[0]) if d <= 5 else None)
    g["c4"] = {"config": W.C4["name"], "first300": [{"t": t, "depth": d, "seed": s, "pos": hexpos(r), "perft": n}
                                                  for (t, d, s), r, n in zip(reqs, rrows, rc)]}
    print("c4 done %.1fs" % (time.time() - t0), flush=True)

    with open(os.path.join(HERE, "workloads.json"), "w") as f:
        json.dump(g, f, indent=0)
    print("wrote workloads.json in %.1fs" % (time.time() - t0))


if __name__ == "__main__":
    main()
