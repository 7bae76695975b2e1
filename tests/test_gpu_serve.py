"""Serving: continuous batching on the GPU.  Every result must equal the same
request's standalone count (oracle / golden), whether it ran alone or in a mix
with requests of other depths arriving while it was on the device."""
import json
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _drain(eng, n, timeout=120.0):
    out = {}
    t0 = time.time()
    while len(out) < n:
        for r in eng.poll():
            out[r["tag"]] = r
        if time.time() - t0 > timeout:
            raise TimeoutError("only %d of %d results" % (len(out), n))
        time.sleep(0.001)
    return out


@pytest.mark.parametrize("tt", [0, 256 << 20])
def test_serving_alone_equals_mixed(pg, oracle, tt):
    eng = pg.Engine(backend="ks", arena_bytes=2 << 30, tt_bytes=tt, zobrist_seed=11)
    try:
        eng.serve_start(max_jobs=16)
        pos = pg.playouts(1234, 40, 4, 60)
        depths = [1 + (i % 5) for i in range(40)]  # 1..5
        want = [int(oracle.perft_batch(pos[i:i + 1], depths[i], nthreads=1)[0]) for i in range(40)]
        # alone: one request at a time
        alone = {}
        for i in range(0, 40, 8):
            eng.submit(pos[i], depths[i], 1000 + i)
            alone.update(_drain(eng, 1))
        for i in range(0, 40, 8):
            assert alone[1000 + i]["nodes"] == want[i] and alone[1000 + i]["status"] == 0
        # mixed: all 40 at once, then a second wave while the first runs
        for i in range(40):
            eng.submit(pos[i], depths[i], i)
        for i in range(40):
            eng.submit(pos[i], depths[i], 100 + i)
        got = _drain(eng, 80)
        for i in range(40):
            assert got[i]["status"] == 0 and got[i]["nodes"] == want[i], (i, depths[i])
            assert got[100 + i]["nodes"] == want[i]
            assert got[i]["done_s"] >= got[i]["submit_s"]
        if tt:  # depth-5 requests write a level at ply 3: duplicates were found and merged
            assert eng.stats()["dup_nodes"] > 0
        eng.serve_stop()
        # the context is usable for plain perft again after serving
        assert eng.perft(pg.fen_to_pos(pg.START_FEN), 4) == 197281
    finally:
        eng.close()


def test_serving_c4_requests_vs_golden(pg):
    with open(os.path.join(GOLDEN, "workloads.json")) as f:
        reqs = json.load(f)["c4"]["first300"]
    eng = pg.Engine(backend="ks", arena_bytes=8 << 30, tt_bytes=1 << 30, zobrist_seed=0xC4)
    try:
        eng.serve_start(max_jobs=64)
        n = 0
        for i, r in enumerate(reqs):
            if r["perft"] is None:
                continue
            eng.submit(tuple(int(x, 16) for x in r["pos"]), r["depth"], i)
            n += 1
        got = _drain(eng, n)
        for i, r in enumerate(reqs):
            if r["perft"] is not None:
                assert got[i]["nodes"] == r["perft"] and got[i]["status"] == 0, i
        # busy while serving
        with pytest.raises(pg.PerftError) as ei:
            eng.perft(pg.fen_to_pos(pg.START_FEN), 3)
        assert ei.value.status == pg.PF_EBUSY
        eng.serve_stop()
    finally:
        eng.close()


def test_serving_chunked_small_slots(pg, oracle):
    """1 MiB per job forces depth-first chunking inside every job"""
    eng = pg.Engine(backend="ks", arena_bytes=8 << 20, tt_bytes=8 << 20, zobrist_seed=5)
    try:
        eng.serve_start(max_jobs=8, job_bytes=1 << 20)
        pos = pg.playouts(999, 8, 10, 40)
        want = oracle.perft_batch(pos, 5, nthreads=8)
        for i in range(8):
            eng.submit(pos[i], 5, i)
        got = _drain(eng, 8)
        assert [got[i]["nodes"] for i in range(8)] == [int(x) for x in want]
        eng.serve_stop()
    finally:
        eng.close()


def test_serving_chunks_by_block_sums(pg, oracle):
    """16 MiB per job: the ply-4 levels of depth-6 requests do not fit, and the
    jobs chunk their ply-3 parents by exact block sums of their move counts"""
    eng = pg.Engine(backend="ks", arena_bytes=64 << 20, tt_bytes=64 << 20, zobrist_seed=9)
    try:
        eng.serve_start(max_jobs=4, job_bytes=16 << 20)
        pos = pg.playouts(2024, 4, 12, 30)
        want = oracle.perft_batch(pos, 6, nthreads=8)
        for i in range(4):
            eng.submit(pos[i], 6, i)
        got = _drain(eng, 4, timeout=300.0)
        assert [got[i]["nodes"] for i in range(4)] == [int(x) for x in want]
        assert all(got[i]["status"] == 0 for i in range(4))
        eng.serve_stop()
    finally:
        eng.close()


def test_serving_no_table_leaf_parent_chunks(pg):
    """No table (so no keys and no duplicate tables): 64 MiB per job, and the
    ply-4 levels of depth-6 requests (about 1.5e6 nodes x 80 B each) do not
    fit, so the ply-3 frames chunk right above the leaf pass; checked against
    the oracle's perft 6 of the first C5 positions"""
    with open(os.path.join(GOLDEN, "workloads.json")) as f:
        want = json.load(f)["c5"]["perft_first8_d6"][:4]
    from perft_gpu_amd import workloads as W
    pos = W.positions(W.C5)[:4]
    eng = pg.Engine(backend="ks", arena_bytes=256 << 20, tt_bytes=0)
    try:
        eng.serve_start(max_jobs=4, job_bytes=64 << 20)
        for i in range(4):
            eng.submit(pos[i], 6, i)
        got = _drain(eng, 4, timeout=300.0)
        assert [got[i]["status"] for i in range(4)] == [0] * 4
        assert [got[i]["nodes"] for i in range(4)] == want
        eng.serve_stop()
    finally:
        eng.close()


def test_serving_job_grows_into_free_memory(pg):
    """a 256 MiB arena caps each job's predicted region at 64 MiB, where the
    ply-4 level of a depth-6 request (about 1.5e6 nodes) does not fit.  Alone
    (nothing waiting for admission) the job takes the free memory after its
    region and runs the level whole, in fewer steps than the same request in a
    fixed 64 MiB region, which chunks it; both equal the oracle's count"""
    with open(os.path.join(GOLDEN, "workloads.json")) as f:
        want = json.load(f)["c5"]["perft_first8_d6"][:2]
    from perft_gpu_amd import workloads as W
    pos = W.positions(W.C5)[:2]
    steps = {}
    for name, job_bytes in (("grows", 0), ("fixed", 64 << 20)):
        eng = pg.Engine(backend="ks", arena_bytes=256 << 20, tt_bytes=64 << 20, zobrist_seed=5)
        try:
            eng.serve_start(max_jobs=4, job_bytes=job_bytes)
            got = []
            for i in range(2):  # one at a time, so that nothing waits while it runs
                eng.submit(pos[i], 6, i)
                r = _drain(eng, 1, timeout=300.0)[i]
                assert r["status"] == 0
                got.append(r["nodes"])
            steps[name] = eng.stats()["levels_run"]
            eng.serve_stop()
        finally:
            eng.close()
        assert got == want, name
    assert steps["grows"] < steps["fixed"], steps
