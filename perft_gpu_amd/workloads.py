"""The five benchmark workloads as seeded, reproducible definitions.

Positions come from pf_playouts (seeded xorshift64 random legal playouts from
the initial position; see include/perft_gpu_amd.h for the exact procedure).

  C1  initial position, depth 5, Kogge-Stone, no TT (CPU path of the reference)
  C2  256 playouts of 8..40 plies (mate/stalemate discarded), depth 5 each,
      one persistent launch, no TT, four slider backends compared
  C3  one 20..30-ply playout with >= 35 legal moves, depth 8, 64 GiB TT,
      depth-first chunking when a level exceeds memory
  C4  serving: Poisson arrivals at 50 req/s for 120 s, playouts of 10..60
      plies, depth drawn from {3:.30, 4:.30, 5:.25, 6:.12, 7:.03}, 32 GiB TT
  C5  4096 playouts of 15..50 plies with >= 35 legal moves and no check,
      depth 6 each, 128 GiB TT
"""
import math

from . import START_FEN, fen_to_pos, playouts

C1 = dict(name="startpos_d5_ks", fen=START_FEN, depth=5, backend="ks")
C2 = dict(name="c2_256x_playout8-40_d5", seed=0xC2C2_0256, n=256, min_plies=8, max_plies=40, min_moves=0,
          no_check=False, depth=5)
C3 = dict(name="c3_middlegame_d8_tt64g", seed=0xC3C3_0008, n=1, min_plies=20, max_plies=30, min_moves=35,
          no_check=False, depth=8, tt_bytes=64 << 30)
C4 = dict(name="c4_serving_poisson50_120s", seed=0xC4C4_0050, rate=50.0, seconds=120.0, min_plies=10,
          max_plies=60, depths=(3, 4, 5, 6, 7), probs=(0.30, 0.30, 0.25, 0.12, 0.03), tt_bytes=32 << 30)
C5 = dict(name="c5_4096x_highmob_d6_tt128g", seed=0xC5C5_4096, n=4096, min_plies=15, max_plies=50,
          min_moves=35, no_check=True, depth=6, tt_bytes=128 << 30)


def positions(cfg):
    if "fen" in cfg:
        import numpy as np
        from . import POS_DTYPE
        arr = np.zeros(1, dtype=POS_DTYPE)
        arr[0] = fen_to_pos(cfg["fen"]).as_tuple()
        return arr
    return playouts(cfg["seed"], cfg["n"], cfg["min_plies"], cfg["max_plies"], cfg["min_moves"], cfg["no_check"])


def _xs(x):
    m = (1 << 64) - 1
    x ^= (x << 13) & m
    x ^= x >> 7
    x ^= (x << 17) & m
    return x


def serving_requests(cfg=C4, limit=None):
    """[(arrival_s, depth, position_seed)] for the serving workload.

    Inter-arrival gaps are exponential with rate `rate` (inverse CDF of a
    53-bit uniform from a xorshift64 stream); depth by inverse CDF over
    `probs`; the request's position is playouts(position_seed, 1, 10, 60)."""
    x = cfg["seed"]
    t = 0.0
    out = []
    cum = []
    c = 0.0
    for p in cfg["probs"]:
        c += p
        cum.append(c)
    while True:
        x = _xs(x)
        u = ((x >> 11) + 0.5) / float(1 << 53)
        t += -math.log(u) / cfg["rate"]
        if t >= cfg["seconds"]:
            break
        x = _xs(x)
        v = (x >> 11) / float(1 << 53)
        depth = cfg["depths"][-1]
        for d, cp in zip(cfg["depths"], cum):
            if v < cp:
                depth = d
                break
        x = _xs(x)
        out.append((t, depth, x | 1))
        if limit and len(out) >= limit:
            break
    return out


def request_position(pos_seed, cfg=C4):
    return playouts(pos_seed, 1, cfg["min_plies"], cfg["max_plies"], 0, False)[0]

# Algorithmic op model of the C2 workload: int32 lane-ops of the breadth-first
# reference algorithm with Kogge-Stone sliders (generate at interior nodes,
# make+count at every node below the root; 64-bit and/or/xor/not/shift/add/
# sub/select = 2, multiply = 4, popcount = 2, ctz = 3, table loads = 0),
# summed over the whole C2 tree by the instrumented CPU restatement
# (tests/golden/make_golden.py): 416,777,397,783 ops / 8,988,346,871 leaves.
C2_OPS_PER_LEAF = 416777397783 / 8988346871
C2_LEAVES = 8988346871
# The same op total per node made below the root (plies 1-4 of the C2 tree:
# 8,001 + 249,956 + 8,164,767 + 265,225,172 = 273,647,896 nodes, each made and
# counted, the interior ones also generated): the lane-op cost of one made
# and counted node, used to price the counting work of table runs.
C2_MADE_NODES = 8001 + 249956 + 8164767 + 265225172
C2_OPS_PER_MADE_NODE = 416777397783 / C2_MADE_NODES
# What the hardware issues for the same step with the Kogge-Stone kernels
# (rocprofv3 --pmc SQ_INSTS_VALU on one C2 step: 3.3140e9 wave-instructions in
# the leaf level, 1.1059e8 in the three write levels), per leaf, as lane-
# instructions; and the most wave-instructions per SIMD-cycle the leaf loop's
# instruction mix can issue: 509 of its 841 VALU instructions issue every
# ~2.4 cycles (two-input logic, bitop3, add, moves), 332 every ~4.2 (shifts,
# popcounts, selects, compares, perms), measured per opcode with all waves
# resident, i.e. 841 instructions per 2,616 cycles.
C2_VALU_LANE_INSTR_PER_LEAF = (3.3140e9 + 1.1059e8) * 64 / 8988346871
C2_LOOP_ISSUE_LIMIT = 841 / 2616
