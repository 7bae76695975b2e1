"""CPU-side checks of the product: the C ABI library loads and exports every
declared symbol, the host helpers (FEN, playouts, UCI moves) match the oracle,
and the shared host/device move generator (movegen.h), built on the CPU by a
test-only harness, matches the oracle on known answers and random trees."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, build_harness
from oracle_lib import Oracle

HEADER = os.path.join(ROOT, "include", "perft_gpu_amd.h")


def test_header_symbols_exported(pg):
    lib = pg.load_library()
    text = open(HEADER).read()
    declared = sorted(set(re.findall(r"\b(pf_[a-z0-9_]+)\s*\(", text)))
    assert declared, "no declarations found"
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert set(pg.EXPORTS) <= set(declared)


def test_init_without_gpu_fails_loudly(pg):
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(pg.PerftError) as ei:
        pg.Engine()
    assert ei.value.status in (pg.PF_EDEVICE, pg.PF_EINVAL)


def test_fen_parse_matches_oracle(pg, oracle):
    fens = ["rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1",
            "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq -",
            "8/8/1k6/2b5/2pP4/8/5K2/8 b - d3 0 1",
            "r2q1rk1/pP1p2pp/Q4n2/bbp1p3/Np6/1B3NBn/pPPP1PPP/R3K2R b KQ - 0 1"]
    for row in oracle.playouts(3, 30, 1, 100):
        fens.append(oracle.fen(oracle.from48(row)))
    for f in fens:
        assert pg.fen_to_pos(f).as_tuple() == oracle.pos(f).as_tuple(), f
        assert pg.pos_to_fen(pg.fen_to_pos(f)) == oracle.fen(oracle.pos(f))


def test_bad_fen_rejected(pg):
    for bad in ["", "8/8/8/8/8/8/8/8 w - - 0 1", "rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNX w KQkq - 0 1",
                "P7/8/8/8/8/8/8/K6k w - - 0 1"]:
        with pytest.raises(pg.PerftError) as ei:
            pg.fen_to_pos(bad)
        assert ei.value.status == pg.PF_EINVAL


@pytest.mark.parametrize("params", [(1, 64, 8, 40, 0, False), (12345, 48, 15, 50, 35, True),
                                    (99, 40, 10, 60, 0, False), (2, 16, 20, 30, 35, False)])
def test_playouts_match_oracle(pg, oracle, params):
    seed, n, lo, hi, mm, nc = params
    a = pg.playouts(seed, n, lo, hi, mm, nc)
    b = oracle.playouts(seed, n, lo, hi, mm, nc)
    assert len(a) == len(b) == n
    for i in range(n):
        assert tuple(int(x) for x in a[i]) == tuple(int(x) for x in b[i]), i


def test_host_count_matches_oracle(pg, oracle):
    lib = pg.load_library()
    for row in pg.playouts(8, 200, 1, 120):
        p = pg.Pos48(*[int(x) for x in row])
        assert lib.pf_count_moves_host(ctypes.byref(p)) == oracle.count(oracle.from48(row))


def test_apply_uci_and_move_text(pg, oracle):
    p = pg.fen_to_pos(pg.START_FEN)
    for mv in ["e2e4", "c7c5", "g1f3", "d7d6", "d2d4", "c5d4", "f3d4", "g8f6", "b1c3", "a7a6"]:
        p = pg.apply_uci(p, mv)
    assert pg.pos_to_fen(p).startswith("rnbqkb1r/1p2pppp/p2p1n2/8/3NP3/2N5/PPP2PPP/R1BQKB1R w KQkq")
    # promotion text and castling move text
    q = pg.fen_to_pos("r3k2r/1P6/8/8/8/8/8/R3K2R w KQkq - 0 1")
    q2 = pg.apply_uci(q, "b7a8q")
    assert pg.pos_to_fen(q2).startswith("Q3k2r/8")
    q3 = pg.apply_uci(q, "e1g1")
    assert pg.pos_to_fen(q3).startswith("r3k2r/1P6/8/8/8/8/8/R4RK1 b kq")
    with pytest.raises(pg.PerftError):
        pg.apply_uci(q, "e1e3")


# ---- product move generator on the CPU (test harness) ----------------------

@pytest.fixture(scope="module")
def mg(pg):
    lib = ctypes.CDLL(build_harness())
    lib.mg_init.restype = ctypes.c_int
    assert lib.mg_init() == 0
    lib.mg_perft.restype = ctypes.c_uint64
    lib.mg_perft.argtypes = [ctypes.POINTER(pg.Pos48), ctypes.c_int, ctypes.c_int]
    lib.mg_gen_count_agree.argtypes = [ctypes.POINTER(pg.Pos48), ctypes.c_int]
    lib.mg_attacks.restype = ctypes.c_uint64
    lib.mg_attacks.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
    lib.mg_children.argtypes = [ctypes.POINTER(pg.Pos48), ctypes.c_void_p]
    return lib


@pytest.mark.parametrize("backend", [0, 1, 2, 3])
def test_movegen_known_answers(pg, mg, backend):
    from test_oracle import known_answers
    for name, fen, d, want in known_answers(3_000_000):
        p = pg.fen_to_pos(fen)
        assert mg.mg_perft(ctypes.byref(p), d, backend) == want, (backend, name, d)


def test_movegen_random_trees_vs_oracle(pg, oracle, mg):
    rows = pg.playouts(31337, 150, 4, 90)
    for row in rows:
        p = pg.Pos48(*[int(x) for x in row])
        for d in (1, 2, 3):
            assert mg.mg_perft(ctypes.byref(p), d, 0) == oracle.perft_pos48(row, d)
        assert mg.mg_gen_count_agree(ctypes.byref(p), 3) == 0


def test_movegen_children_match_oracle(pg, oracle, mg):
    """make_move + colour flip reproduce the reference's child boards"""
    rows = pg.playouts(2718, 80, 1, 90)
    for row in rows:
        p = pg.Pos48(*[int(x) for x in row])
        out = np.zeros(256, dtype=pg.POS_DTYPE)
        n = mg.mg_children(ctypes.byref(p), out.ctypes.data_as(ctypes.c_void_p))
        o = oracle.from48(row)
        want = sorted(tuple(oracle.make(o, m, 1).as_tuple()) for m in oracle.moves(o))
        got = sorted(tuple(int(x) for x in out[i]) for i in range(n))
        assert got == want


def test_movegen_slider_tables_vs_oracle(oracle, mg):
    rng = np.random.default_rng(3)
    occs = [int(x) for x in rng.integers(0, 2 ** 63, size=64, dtype=np.int64)] + [0, (1 << 64) - 1]
    for sq in range(64):
        for occ in occs:
            occ &= (1 << 64) - 1
            for b in (0, 1, 2, 3):
                assert mg.mg_attacks(b, 1, sq, occ) == oracle.rook_attacks(0, sq, occ)
                assert mg.mg_attacks(b, 0, sq, occ) == oracle.bishop_attacks(0, sq, occ)


def test_workload_positions_match_golden(pg):
    """the product's seeded workloads reproduce the oracle-generated fixtures"""
    import json
    from perft_gpu_amd import workloads as W
    with open(os.path.join(ROOT, "tests", "golden", "workloads.json")) as f:
        g = json.load(f)
    c2 = W.positions(W.C2)
    assert [["%016x" % int(x) for x in r] for r in c2] == g["c2"]["positions"]
    c3 = W.positions(W.C3)
    assert [["%016x" % int(x) for x in r] for r in c3] == g["c3"]["positions"]
    c5 = W.positions(W.C5)
    assert len(c5) == 4096 and [["%016x" % int(x) for x in r] for r in c5[:64]] == g["c5"]["positions_first64"]
    reqs = W.serving_requests(W.C4, limit=300)
    for (t, d, s), want in zip(reqs, g["c4"]["first300"]):
        assert (t, d, s) == (want["t"], want["depth"], want["seed"])
        assert ["%016x" % int(x) for x in W.request_position(s)] == want["pos"]
    assert sum(g["c2"]["perft"]["5"]) == W.C2_LEAVES


def test_zobrist_full_keys_match_oracle(pg, oracle, mg):
    mg.mg_zkey_full.restype = ctypes.c_uint64
    mg.mg_zkey_full.argtypes = [ctypes.POINTER(pg.Pos48), ctypes.c_uint64, ctypes.c_int]
    salt = 0xD1B54A32D192ED03
    for seed in (1, 0xC3C3):
        t1, t2 = oracle.zobrist_table(seed), oracle.zobrist_table(seed ^ salt)
        for row in pg.playouts(seed, 60, 1, 90):
            p = pg.Pos48(*[int(x) for x in row])
            o = oracle.from48(row)
            assert mg.mg_zkey_full(ctypes.byref(p), seed, 0) == oracle.zobrist_key(o, t1)
            assert mg.mg_zkey_full(ctypes.byref(p), seed, 1) == oracle.zobrist_key(o, t2)


def test_zobrist_incremental_keys(pg, mg):
    """the device's incremental key update equals a full recomputation"""
    mg.mg_zkey_children_check.argtypes = [ctypes.POINTER(pg.Pos48), ctypes.c_uint64]
    rows = list(pg.playouts(4, 300, 1, 110))
    rows += [tuple(pg.fen_to_pos(f).as_tuple()) for f in
             ["r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 0 1",
              "r3k2r/Pppp1ppp/1b3nbN/nP6/BBP1P3/q4N2/Pp1P2PP/R2Q1RK1 w kq - 0 1",
              "8/8/1k6/2b5/2pP4/8/5K2/8 b - d3 0 1", "r3k2r/1P6/8/8/8/8/6p1/R3K2R b KQkq - 0 1"]]
    for row in rows:
        p = pg.Pos48(*[int(x) for x in row])
        assert mg.mg_zkey_children_check(ctypes.byref(p), 99) == 0, pg.pos_to_fen(p)


def test_three_input_logic_helpers(mg):
    """board.h's bitop3 helpers (host form: the same truth table the GPU
    instruction takes) equal the two-input expressions they replace"""
    mg.mg_bitop3_check.argtypes = [ctypes.c_uint64, ctypes.c_int]
    assert mg.mg_bitop3_check(20250101, 20000) == 0
