"""Pins for the oracle (CPU restatement of the reference path).

Each check is chosen so that a plausible slip in the restatement -- a dropped
term, a wrong shift or square index, a swapped colour -- fails it:
  * published perft known answers, incl. the en-passant / castling /
    promotion / stalemate edge cases, on all four slider backends;
  * an independent mailbox generator (oracle/naive.c) on random positions;
  * brute force over every (square, occupancy subset) for every slider table,
    against a pure-Python ray walk;
  * closed forms: xorshift64 stream, Zobrist table = that stream, FEN round trip;
  * structural invariants: count == #generated moves == #generated boards,
    makeMove(generateMoves) == generateBoards, sum of divide == perft.
"""
import json
import os
import random

import numpy as np
import pytest

from oracle_lib import BYTE, FANCY, KS, PLAIN, Oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MASK_STATE = 0x00FFFFFFFFFFFF00  # pawn squares; state bytes excluded


def known_answers(limit):
    with open(os.path.join(GOLDEN, "cpw_perft.json")) as f:
        positions = json.load(f)["positions"]
    out = []
    for e in positions:
        for d, want in enumerate(e.get("perft", []), start=1):
            out.append((e["name"], e["fen"], d, want))
        for d, want in e.get("perft_at", {}).items():
            out.append((e["name"], e["fen"], int(d), want))
    return [c for c in out if c[3] <= limit]


@pytest.mark.parametrize("backend", [KS, PLAIN, FANCY, BYTE])
def test_known_answers_all_backends(oracle, backend):
    prev = oracle.set_backend(backend)
    try:
        for name, fen, d, want in known_answers(20_000_000 if backend == FANCY else 5_000_000):
            assert oracle.perft_fen(fen, d) == want, (backend, name, d)
    finally:
        oracle.set_backend(prev)


def test_oracle_vs_independent_mailbox_generator(oracle):
    pos = oracle.playouts(4242, 60, 6, 70)
    assert len(pos) == 60
    for row in pos:
        p = oracle.from48(row)
        for d in (1, 2, 3):
            assert oracle.perft(p, d) == oracle.naive_perft(p, d), (oracle.fen(p), d)


def test_mailbox_generator_known_answers(oracle):
    # the checker itself is pinned by the published answers
    for name, fen, d, want in known_answers(300_000):
        assert oracle.naive_perft(oracle.pos(fen), d) == want, (name, d)


def _ray_walk(sq, occ, dirs):
    att = 0
    f0, r0 = sq & 7, sq >> 3
    for df, dr in dirs:
        f, r = f0 + df, r0 + dr
        while 0 <= f < 8 and 0 <= r < 8:
            att |= 1 << (r * 8 + f)
            if occ >> (r * 8 + f) & 1:
                break
            f += df
            r += dr
    return att


ROOK_DIRS = [(1, 0), (-1, 0), (0, 1), (0, -1)]
BISHOP_DIRS = [(1, 1), (1, -1), (-1, 1), (-1, -1)]


def _subsets(mask):
    bits = [i for i in range(64) if mask >> i & 1]
    for k in range(1 << len(bits)):
        o = 0
        for j, b in enumerate(bits):
            if k >> j & 1:
                o |= 1 << b
        yield o


def test_slider_tables_exhaustive(oracle):
    """every (square, relevant-occupancy subset): all backends == ray walk"""
    n_rook = n_bishop = 0
    for sq in range(64):
        rm = oracle.lib.orc_rook_mask(sq)
        for occ in _subsets(rm):
            want = _ray_walk(sq, occ, ROOK_DIRS)
            for b in (KS, PLAIN, FANCY, BYTE):
                assert oracle.rook_attacks(b, sq, occ) == want, (b, sq, occ)
            n_rook += 1
        bm = oracle.lib.orc_bishop_mask(sq)
        for occ in _subsets(bm):
            want = _ray_walk(sq, occ, BISHOP_DIRS)
            for b in (KS, PLAIN, FANCY, BYTE):
                assert oracle.bishop_attacks(b, sq, occ) == want, (b, sq, occ)
            n_bishop += 1
    assert (n_rook, n_bishop) == (102400, 5248)


def test_fancy_table_sizes(oracle):
    import ctypes
    fe, ru, bu = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    oracle.lib.orc_table_stats(ctypes.byref(fe), ctypes.byref(ru), ctypes.byref(bu))
    # GlobalVars.cpp:42, 187-189: 97,264 shared entries, 4,900 / 1,428 unique attack sets
    assert (fe.value, ru.value, bu.value) == (97264, 4900, 1428)


def test_xorshift64_closed_form(oracle):
    import ctypes
    s = ctypes.c_uint64(1)
    got = [oracle.lib.orc_xorshift64(ctypes.byref(s)) for _ in range(3)]
    assert got == [0x40822041, 0x100041060C011441, 0x9B1E842F6E862629]

    def xs(x):
        m = (1 << 64) - 1
        x ^= (x << 13) & m
        x ^= x >> 7
        x ^= (x << 17) & m
        return x

    t = oracle.zobrist_table(12345)
    x = 12345
    for i in range(781):
        x = xs(x)
        assert t[i] == x


def test_zobrist_key_structure(oracle):
    t = oracle.zobrist_table(7)
    p = oracle.pos("4k3/8/8/8/8/8/8/4K3 w - - 0 1")
    # white king e1 (color 0, type 5, sq 4), black king e8 (1, 5, 60), white to move
    assert oracle.zobrist_key(p, t) == t[(0 * 6 + 5) * 64 + 4] ^ t[(1 * 6 + 5) * 64 + 60] ^ t[780]
    q = oracle.pos("4k3/8/8/8/8/8/8/4K3 b - - 0 1")
    assert oracle.zobrist_key(q, t) == t[(0 * 6 + 5) * 64 + 4] ^ t[(1 * 6 + 5) * 64 + 60]
    r = oracle.pos("r3k3/8/8/8/8/8/8/4K3 b q - 0 1")
    assert oracle.zobrist_key(r, t) == (t[(0 * 6 + 5) * 64 + 4] ^ t[(1 * 6 + 5) * 64 + 60] ^
                                        t[(1 * 6 + 3) * 64 + 56] ^ t[771])


def test_fen_round_trip(oracle):
    for name, fen, _, _ in known_answers(10):
        p = oracle.pos(fen)
        assert oracle.pos(oracle.fen(p)).as_tuple() == p.as_tuple(), name
    for row in oracle.playouts(77, 40, 1, 80):
        p = oracle.from48(row)
        assert oracle.pos(oracle.fen(p)).as_tuple() == p.as_tuple()


def test_fen_quirks(oracle):
    # util.cpp:267-274: only the file of the en-passant square is kept
    a = oracle.pos("4k3/8/8/3pP3/8/8/8/4K3 w - d6 0 1")
    b = oracle.pos("4k3/8/8/3pP3/8/8/8/4K3 w - d3 0 1")
    assert a.as_tuple() == b.as_tuple()
    assert (a.pawns >> 4) & 15 == 4
    # clocks are ignored
    c = oracle.pos("4k3/8/8/3pP3/8/8/8/4K3 w - d6 37 90")
    assert c.as_tuple() == a.as_tuple()


def _norm(p):
    # compare boards without the half-move counter bits (56-62)
    t = list(p.as_tuple())
    t[1] &= ~(0x7F << 56)
    return tuple(t)


def test_generation_invariants(oracle):
    rows = oracle.playouts(99, 80, 1, 90)
    for row in rows:
        p = oracle.from48(row)
        moves = oracle.moves(p)
        kids = oracle.children(p)
        assert oracle.count(p) == len(moves) == len(kids)
        made = sorted(_norm(oracle.make(p, m, 0)) for m in moves)
        assert made == sorted(_norm(k) for k in kids)


def test_divide_sums_to_perft(oracle):
    p = oracle.pos("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 0 1")
    div = oracle.divide(p, 3)
    assert len(div) == 48
    assert sum(c for _, c in div) == 97862


def test_batch_threads_equal_sequential(oracle):
    rows = oracle.playouts(5, 24, 8, 40)
    a = oracle.perft_batch(rows, 3, nthreads=1)
    b = oracle.perft_batch(rows, 3, nthreads=8)
    assert np.array_equal(a, b)
    assert [oracle.perft_pos48(r, 3) for r in rows] == list(map(int, a))


def test_exact_ep_flag_only_when_capturable(oracle):
    p = oracle.pos("4k3/8/8/8/8/8/4P3/4K3 w - - 0 1")
    e2e4 = 12 | (28 << 6) | (1 << 12)
    assert (oracle.make(p, e2e4, 1).pawns >> 4) & 15 == 0
    assert (oracle.make(p, e2e4, 0).pawns >> 4) & 15 == 5
    q = oracle.pos("4k3/8/8/8/3p4/8/4P3/4K3 w - - 0 1")
    assert (oracle.make(q, e2e4, 1).pawns >> 4) & 15 == 5


def test_playout_filters(oracle):
    rows = oracle.playouts(11, 30, 15, 50, min_moves=35, no_check=True)
    assert len(rows) == 30
    for r in rows:
        p = oracle.from48(r)
        assert oracle.count(p) >= 35 and not oracle.in_check(p)


def test_op_model(oracle):
    p = oracle.pos("rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1")
    ops, leaves = oracle.opcount(p, 4)
    assert leaves == 197281
    # one count per depth-1 node amortised over ~22 leaves at this depth
    assert 10 < ops / leaves < 200
    ops5, leaves5 = oracle.opcount(p, 5)
    assert leaves5 == 4865609 and ops5 > ops


def test_op_model_constant_matches_golden():
    """the op-model constant used by bench.py is the one the fixture records"""
    import json
    from perft_gpu_amd import workloads as W
    with open(os.path.join(GOLDEN, "workloads.json")) as f:
        om = json.load(f)["c2"]["op_model"]
    assert om["leaves"] == W.C2_LEAVES
    assert abs(om["int32_lane_ops"] / om["leaves"] - W.C2_OPS_PER_LEAF) < 1e-12
