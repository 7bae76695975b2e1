"""Configuration-scale parity on the GPU: the configured table sizes (C3 64 GiB,
C5 128 GiB, C4 32 GiB), the planned table passes (one captured graph per
chunk) against the host-driven levels, deep C4 requests served in a mix, the
byte-lookup backend with a table, and the UCI front end -- all against the
oracle's golden counts (tests/golden/workloads.json and deep.json)."""
import json
import os
import re
import subprocess
import time

import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _gold():
    with open(os.path.join(GOLDEN, "workloads.json")) as f:
        g = json.load(f)
    deep = os.path.join(GOLDEN, "deep.json")
    if os.path.exists(deep):
        with open(deep) as f:
            g["deep"] = json.load(f)
    return g


def test_c3_perft7_with_the_64g_table(pg):
    """C3's position at perft 7 with its configured 64 GiB table: planned
    passes on (the default) and off, each on a fresh table"""
    from perft_gpu_amd import workloads as W
    want = _gold()["c3"]["perft"]["7"][0]
    pos = W.positions(W.C3)
    for graphs in (True, False):
        print("c3 perft 7, 64 GiB table, graphs %s" % graphs, flush=True)
        eng = pg.Engine(backend="ks", arena_bytes=16 << 30, tt_bytes=W.C3["tt_bytes"], zobrist_seed=W.C3["seed"])
        try:
            eng.set_level_graphs(graphs)
            assert int(eng.perft_batch(pos, 7)[0]) == want
            st = eng.stats()
            assert (st["graph_passes"] > 0) == graphs
        finally:
            eng.close()


def test_c5_roots_perft6_with_the_128g_table_chunked(pg):
    """the first 64 C5 positions at perft 6 with the configured 128 GiB table
    and a 3 GiB arena, so that the roots run in several chunks: planned passes
    against the host-driven levels, both against the oracle"""
    from perft_gpu_amd import workloads as W
    g = _gold()
    want = g["deep"]["c5"]["perft6"][:64] if "deep" in g and "c5" in g["deep"] else None
    pos = W.positions(W.C5)[:64]
    got = {}
    for graphs in (True, False):
        print("c5 64 roots perft 6, 128 GiB table, 3 GiB arena, graphs %s" % graphs, flush=True)
        eng = pg.Engine(backend="ks", arena_bytes=3 << 30, tt_bytes=W.C5["tt_bytes"], zobrist_seed=W.C5["seed"])
        try:
            eng.set_level_graphs(graphs)
            got[graphs] = [int(x) for x in eng.perft_batch(pos, 6)]
            st = eng.stats()
            if graphs:
                assert st["graph_passes"] > 1
            else:
                assert st["graph_passes"] == 0 and st["chunks"] > 0
        finally:
            eng.close()
    assert got[True] == got[False]
    assert got[True][:8] == g["c5"]["perft_first8_d6"]
    if want is not None:
        assert got[True] == want


def test_serving_deep_c4_requests_in_a_mix(pg):
    """depth-6 and depth-7 requests of the C4 stream, served together with
    shallow ones on the configured 32 GiB table: every count equals the
    oracle's"""
    g = _gold()
    if "deep" not in g or "requests" not in g["deep"].get("c4", {}):
        pytest.skip("deep C4 goldens not generated")
    from perft_gpu_amd import workloads as W
    reqs = g["deep"]["c4"]["requests"]
    d7 = [i for i, r in enumerate(reqs) if r[1] == 7][:12]
    d6 = [i for i, r in enumerate(reqs) if r[1] == 6][:40]
    shallow = [i for i, r in enumerate(reqs) if r[1] <= 5][:80]
    pick = sorted(d7 + d6 + shallow)
    eng = pg.Engine(backend="ks", arena_bytes=48 << 30, tt_bytes=W.C4["tt_bytes"], zobrist_seed=W.C4["seed"])
    try:
        eng.serve_start(max_jobs=256)
        for i in pick:
            eng.submit(W.request_position(reqs[i][2]), reqs[i][1], i)
        out = {}
        t0 = time.time()
        while len(out) < len(pick):
            for r in eng.poll():
                out[r["tag"]] = r
            assert time.time() - t0 < 300, "only %d of %d results" % (len(out), len(pick))
            time.sleep(0.001)
        eng.serve_stop()
    finally:
        eng.close()
    bad = [(i, reqs[i][1], out[i]["nodes"], reqs[i][3]) for i in pick
           if out[i]["status"] != 0 or out[i]["nodes"] != reqs[i][3]]
    assert not bad, bad[:5]


def test_byte_backend_with_a_table(pg):
    """the byte-lookup backend on the table path (its tables read through the
    caches there): batched and plain counts equal the golden ones"""
    gold = _gold()["c2"]["perft"]
    from perft_gpu_amd import workloads as W
    pos = W.positions(W.C2)[:48]
    eng = pg.Engine(backend="byte", arena_bytes=4 << 30, tt_bytes=512 << 20, zobrist_seed=3)
    try:
        for d in (3, 5):
            assert [int(x) for x in eng.perft_batch(pos, d)] == gold[str(d)][:48]
        assert eng.stats()["tt_probes"] > 0
        assert int(eng.perft(pg.Pos48(*[int(x) for x in pos[5]]), 5)) == gold["5"][5]
    finally:
        eng.close()


def test_uci_go_perft_divides_like_the_oracle(pg, oracle):
    """perft_gpu uci: position startpos moves ..., position fen ..., go perft;
    every root move's count and the total against the oracle's divide"""
    cli = pg.CLI_PATH
    script = ("uci\nisready\nposition startpos moves e2e4 e7e5 g1f3\ngo perft 3\n"
              "position fen r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 0 1 moves e1g1\n"
              "go perft 2\nquit\n")
    out = subprocess.run([cli, "uci"], input=script, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    assert "uciok" in lines and "readyok" in lines

    def blocks():
        cur = {}
        for ln in lines:
            if ln.startswith("Nodes searched:"):
                yield cur, int(ln.split(":")[1])
                cur = {}
                continue
            m = re.match(r"^([a-h][1-8][a-h][1-8][nbrq]?): (\d+)$", ln)
            if m:
                cur[m.group(1)] = int(m.group(2))
    got = list(blocks())
    assert len(got) == 2

    def want_of(fen, depth):
        o = oracle.pos(fen)
        w = {}
        for m, c in oracle.divide(o, depth):
            w[pg.move_to_uci(m & 0x0FFF | ((m >> 12) & 0xB) << 12 if (m >> 12) & 8 else m & 0x0FFF)] = c
        return w
    w1 = want_of("rnbqkbnr/pppp1ppp/8/4p3/4P3/5N2/PPPP1PPP/RNBQKB1R b KQkq - 1 2", 3)
    w2 = want_of("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R4RK1 b kq - 1 1", 2)
    assert got[0][0] == w1 and got[0][1] == sum(w1.values())
    assert got[1][0] == w2 and got[1][1] == sum(w2.values())


def test_ep_flagging_and_table_layout_keep_counts(pg):
    """en passant flagged after every double push (the reference CPU
    generator's rule) and the per-depth table layout change keys and where
    entries go, never a count; exact flagging merges more transpositions, so
    it writes no more distinct nodes"""
    from perft_gpu_amd import workloads as W
    g = _gold()
    c3 = W.positions(W.C3)
    c2 = W.positions(W.C2)[:32]
    dups = {}
    for exact in (True, False):
        for regions in (0, 2):
            print("c3 perft 6: exact ep %s, table regions 2^%d" % (exact, regions), flush=True)
            eng = pg.Engine(backend="ks", arena_bytes=4 << 30, tt_bytes=1 << 30, zobrist_seed=W.C3["seed"])
            try:
                eng.set_exact_ep(exact)
                eng.set_tt_regions(regions)
                assert int(eng.perft_batch(c3, 6)[0]) == g["c3"]["perft"]["6"][0]
                st = eng.stats()
                dups[(exact, regions)] = (st["frontier_nodes"], st["dup_nodes"])
                # a second batch over the same roots is answered from the table
                assert [int(x) for x in eng.perft_batch(c2, 5)] == g["c2"]["perft"]["5"][:32]
                eng.reset_stats()
                assert [int(x) for x in eng.perft_batch(c2, 5)] == g["c2"]["perft"]["5"][:32]
                st = eng.stats()
                assert st["tt_hits"] == 32 and st["tt_probes"] == 32
                if regions:
                    with pytest.raises(pg.PerftError) as ei:
                        eng.tt_save("/tmp/never_written.pftt")
                    assert ei.value.status == pg.PF_EINVAL
            finally:
                eng.close()
    assert dups[(True, 0)][0] <= dups[(False, 0)][0]
    assert dups[(True, 0)][1] > 0 and dups[(False, 0)][1] > 0
    assert dups[(True, 0)] == dups[(True, 2)]  # the layout changes no level's contents
