"""Transposition table on the GPU: TT-on counts equal TT-off / golden counts
(hits must never change a count), repeated requests are answered from the
table, and chunking under a tiny arena does not change counts."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden():
    with open(os.path.join(GOLDEN, "workloads.json")) as f:
        return json.load(f)


def _rows(pg, hexrows):
    arr = np.zeros(len(hexrows), dtype=pg.POS_DTYPE)
    for i, r in enumerate(hexrows):
        arr[i] = tuple(int(x, 16) for x in r)
    return arr


@pytest.fixture(scope="module")
def tt_engine(pg):
    eng = pg.Engine(backend="ks", arena_bytes=8 << 30, tt_bytes=1 << 30, zobrist_seed=0xC3C3)
    yield eng
    eng.close()


def test_tt_known_answers(pg, tt_engine):
    from test_gpu_parity import cpw_cases
    for fen, d, want in cpw_cases(20_000_000):
        assert tt_engine.perft(pg.fen_to_pos(fen), d) == want, (fen, d)
    assert tt_engine.stats()["tt_hits"] > 0


def test_tt_c2_vs_golden(pg, tt_engine):
    g = _golden()["c2"]
    pos = _rows(pg, g["positions"])
    for d in (3, 4, 5):
        got = tt_engine.perft_batch(pos, d)
        assert [int(x) for x in got] == g["perft"][str(d)], d
    # second pass: every root is answered from the table
    tt_engine.reset_stats()
    got = tt_engine.perft_batch(pos, 5)
    assert [int(x) for x in got] == g["perft"]["5"]


def test_tt_c3_position_vs_golden(pg, tt_engine):
    g = _golden()["c3"]
    pos = _rows(pg, g["positions"])
    for d in range(3, 7):
        assert int(tt_engine.perft_batch(pos, d)[0]) == g["perft"][str(d)][0], d


def test_tt_c5_subset_vs_golden(pg, tt_engine):
    g = _golden()["c5"]
    pos = _rows(pg, g["positions_first64"])
    assert [int(x) for x in tt_engine.perft_batch(pos[:8], 6)] == g["perft_first8_d6"]
    assert [int(x) for x in tt_engine.perft_batch(pos, 5)] == g["perft_first64"]["5"]


def test_tt_on_equals_tt_off_random(pg, oracle, tt_engine, engine):
    pos = pg.playouts(55, 64, 6, 70)
    for d in (3, 4, 5):
        a = tt_engine.perft_batch(pos, d)
        b = engine.perft_batch(pos, d, mode=pg.MODE_LEVELS)
        assert np.array_equal(a, b), d
    assert np.array_equal(a, oracle.perft_batch(pos, 5, nthreads=8))


def test_tt_chunked_small_arena(pg, oracle):
    eng = pg.Engine(backend="ks", arena_bytes=6 << 20, tt_bytes=64 << 20, zobrist_seed=7)
    try:
        pos = pg.playouts(321, 12, 10, 40)
        want = oracle.perft_batch(pos, 5, nthreads=8)
        assert np.array_equal(eng.perft_batch(pos, 5), want)
        assert eng.stats()["chunks"] > 0
    finally:
        eng.close()


def test_tt_tiny_table_heavy_replacement(pg, oracle):
    """a 64 KiB table forces constant replacement; counts stay exact"""
    eng = pg.Engine(backend="fancy", arena_bytes=1 << 30, tt_bytes=64 << 10, zobrist_seed=3)
    try:
        pos = pg.playouts(8080, 16, 8, 40)
        assert np.array_equal(eng.perft_batch(pos, 5), oracle.perft_batch(pos, 5, nthreads=8))
    finally:
        eng.close()


def test_dedup_on_off_equal_known_answer(pg):
    """in-level duplicate elimination: same count on and off, and it finds
    duplicates (start position: 8,902 nodes at ply 3, far fewer distinct)"""
    start = pg.fen_to_pos(pg.START_FEN)
    for on in (True, False):
        eng = pg.Engine(backend="ks", arena_bytes=2 << 30, tt_bytes=256 << 20, zobrist_seed=11)
        try:
            eng.set_dedup(on)
            assert eng.perft(start, 6) == 119060324
            dups = eng.stats()["dup_nodes"]
            assert (dups > 0) if on else (dups == 0), (on, dups)
        finally:
            eng.close()


def test_dedup_across_roots_of_a_batch(pg, oracle):
    """equal roots in one batch: every node below the first copy's ply 3 is a
    duplicate of a node of the other copy, and both roots get the full count"""
    pos = pg.playouts(4711, 6, 10, 30)
    batch = np.concatenate([pos, pos, pos[::-1]])
    eng = pg.Engine(backend="fancy", arena_bytes=4 << 30, tt_bytes=128 << 20, zobrist_seed=5)
    try:
        got = eng.perft_batch(batch, 5)
        want = oracle.perft_batch(pos, 5, nthreads=8)
        assert np.array_equal(got[:6], want) and np.array_equal(got[6:12], want)
        assert np.array_equal(got[12:], want[::-1])
        assert eng.stats()["dup_nodes"] > 0
    finally:
        eng.close()


def test_dedup_c3_golden_fresh_table(pg):
    g = _golden()["c3"]
    pos = _rows(pg, g["positions"])
    eng = pg.Engine(backend="ks", arena_bytes=16 << 30, tt_bytes=1 << 30, zobrist_seed=0xC3C3)
    try:
        assert int(eng.perft_batch(pos, 6)[0]) == g["perft"]["6"][0]
        assert eng.stats()["dup_nodes"] > 0
    finally:
        eng.close()
