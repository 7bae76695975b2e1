"""GPU parity: the HIP engine (through the C ABI) against the oracle and the
CPW known answers.  Bit-exact integer counts."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def cpw_cases(limit=5_000_000):
    """(fen, depth, count) for every known answer up to `limit` leaves"""
    with open(os.path.join(GOLDEN, "cpw_perft.json")) as f:
        positions = json.load(f)["positions"]
    out = []
    for e in positions:
        for d, want in enumerate(e.get("perft", []), start=1):
            out.append((e["fen"], d, want))
        for d, want in e.get("perft_at", {}).items():
            out.append((e["fen"], int(d), want))
    return [c for c in out if c[2] <= limit]


@pytest.mark.parametrize("backend", ["ks", "plain", "fancy", "byte"])
def test_cpw_known_answers(pg, engine, backend):
    engine.set_backend(backend)
    try:
        for fen, d, want in cpw_cases():
            got = engine.perft(pg.fen_to_pos(fen), d)
            assert got == want, (backend, fen, d, got, want)
    finally:
        engine.set_backend("ks")


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_batch_vs_oracle(pg, oracle, engine, mode):
    arr = pg.playouts(20240601, 96, 8, 60)
    for depth in (1, 2, 3, 4):
        got = engine.perft_batch(arr, depth, mode=mode)
        want = oracle.perft_batch(arr, depth, nthreads=8)
        assert np.array_equal(got, want), (depth, mode)


def _golden():
    with open(os.path.join(GOLDEN, "workloads.json")) as f:
        return json.load(f)


def _rows(pg, hexrows):
    arr = np.zeros(len(hexrows), dtype=pg.POS_DTYPE)
    for i, r in enumerate(hexrows):
        arr[i] = tuple(int(x, 16) for x in r)
    return arr


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_c2_full_size_vs_golden(pg, engine, mode):
    """the full C2 workload (256 positions, perft 5, 9.0e9 leaves)"""
    g = _golden()["c2"]
    pos = _rows(pg, g["positions"])
    for d in (3, 5):
        got = engine.perft_batch(pos, d, mode=mode)
        assert [int(x) for x in got] == g["perft"][str(d)], (d, mode)


@pytest.mark.parametrize("backend", ["plain", "fancy", "byte"])
def test_c2_backends_vs_golden(pg, engine, backend):
    g = _golden()["c2"]
    pos = _rows(pg, g["positions"])
    engine.set_backend(backend)
    try:
        for mode in (pg.MODE_PERSISTENT, pg.MODE_STREAM):
            got = engine.perft_batch(pos, 5, mode=mode)
            assert [int(x) for x in got] == g["perft"]["5"], mode
    finally:
        engine.set_backend("ks")


def test_byte_backend_lds_form_vs_golden(pg, tmp_path):
    """the byte backend's other stream-ordered form, its tables staged in each
    workgroup's LDS, selected through a tune file for both level kinds; the
    tuner may also choose it, and its counts must equal the cache form's"""
    g = _golden()["c2"]
    pos = _rows(pg, g["positions"])
    path = tmp_path / "lds.txt"
    path.write_text("perft_gpu_amd tuning 1\nbackend byte stream_leaf 6 stream_write 6\n")
    eng = pg.Engine(backend="byte", arena_bytes=8 << 30, tune_file=str(path))
    try:
        assert [int(x) for x in eng.perft_batch(pos, 5, mode=pg.MODE_STREAM)] == g["perft"]["5"]
        assert [int(x) for x in eng.perft_batch(pos[:32], 4, mode=pg.MODE_STREAM)] == g["perft"]["4"][:32]
    finally:
        eng.close()
    # the same form number means nothing to the other backends: a tune file
    # naming it for them falls back to their defaults
    path.write_text("perft_gpu_amd tuning 1\nbackend ks stream_leaf 6 stream_write 6\n")
    eng = pg.Engine(backend="ks", arena_bytes=8 << 30, tune_file=str(path))
    try:
        assert [int(x) for x in eng.perft_batch(pos[:32], 4, mode=pg.MODE_STREAM)] == g["perft"]["4"][:32]
    finally:
        eng.close()


def test_c3_c5_positions_vs_golden(pg, engine):
    g = _golden()
    c3 = _rows(pg, g["c3"]["positions"])
    assert [int(engine.perft_batch(c3, d)[0]) for d in range(1, 7)] == [g["c3"]["perft"][str(d)][0] for d in range(1, 7)]
    c5 = _rows(pg, g["c5"]["positions_first64"])
    assert [int(x) for x in engine.perft_batch(c5, 5)] == g["c5"]["perft_first64"]["5"]
    assert [int(x) for x in engine.perft_batch(c5[:8], 6)] == g["c5"]["perft_first8_d6"]


def test_edge_cases(pg, oracle, engine):
    # empty batch, depth 0, mate and stalemate roots, the 218-move position
    assert len(engine.perft_batch(np.zeros(0, dtype=pg.POS_DTYPE), 3)) == 0
    start = pg.fen_to_pos(pg.START_FEN)
    assert engine.perft(start, 0) == 1
    mate = pg.fen_to_pos("rnb1kbnr/pppp1ppp/8/4p3/6Pq/5P2/PPPPP2P/RNBQKBNR w KQkq - 1 3")
    stale = pg.fen_to_pos("7k/5Q2/6K1/8/8/8/8/8 b - - 0 1")
    m218 = pg.fen_to_pos("3Q4/1Q4Q1/4Q3/2Q4R/Q4Q2/3Q4/1Q4Rp/1K1BBNNk w - - 0 1")
    startpos = [1, 20, 400, 8902, 197281]
    for mode in (pg.MODE_PERSISTENT, pg.MODE_STREAM):
        for d in (1, 2, 3, 4):
            got = engine.perft_batch([mate, stale, start], d, mode=mode)
            assert list(map(int, got)) == [0, 0, startpos[d]], (mode, d)
        # a batch whose every root is mate or stalemate: empty levels below the roots
        for d in (2, 3, 4):
            assert list(map(int, engine.perft_batch([mate, stale], d, mode=mode))) == [0, 0], (mode, d)
    assert engine.perft(m218, 1) == 218
    assert engine.perft(m218, 3) == oracle.perft_fen("3Q4/1Q4Q1/4Q3/2Q4R/Q4Q2/3Q4/1Q4Rp/1K1BBNNk w - - 0 1", 3)
    with pytest.raises(pg.PerftError):
        engine.perft(start, 17)


def test_chunking_small_arena(pg, oracle):
    """a 4 MiB arena forces depth-first chunking; counts must not change"""
    eng = pg.Engine(backend="ks", arena_bytes=4 << 20)
    try:
        pos = pg.playouts(777, 16, 10, 40)
        want = oracle.perft_batch(pos, 5, nthreads=8)
        got = eng.perft_batch(pos, 5, mode=pg.MODE_LEVELS)
        assert np.array_equal(got, want)
        assert eng.stats()["chunks"] > 0
        # the single passes cannot chunk: they report ENOMEM, AUTO falls back to levels
        for mode in (pg.MODE_PERSISTENT, pg.MODE_STREAM):
            with pytest.raises(pg.PerftError) as ei:
                eng.perft_batch(pos, 5, mode=mode)
            assert ei.value.status == pg.PF_ENOMEM, mode
        assert np.array_equal(eng.perft_batch(pos, 5, mode=pg.MODE_AUTO), want)
    finally:
        eng.close()


def test_divide(pg, oracle, engine):
    pos = pg.fen_to_pos("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 0 1")
    div = engine.divide(pos, 4)
    assert len(div) == 48 and sum(c for _, c in div) == 4085603
    o = oracle.pos("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 0 1")
    want = {}
    for m, c in oracle.divide(o, 4):
        want[pg.move_to_uci(m & 0x0FFF | ((m >> 12) & 0xB) << 12 if (m >> 12) & 8 else m & 0x0FFF)] = c
    assert dict(div) == want


def test_autotune_keeps_counts_and_round_trips(pg, oracle, tmp_path):
    """the tuner tries every tile shape / workgroups-per-CU choice (each must
    give the same counts), writes its choice, and a context initialised from
    the file gives the same counts"""
    pos = pg.playouts(7070, 32, 8, 40)
    want = oracle.perft_batch(pos, 4, nthreads=8)
    path = str(tmp_path / "tune.txt")
    eng = pg.Engine(backend="ks", arena_bytes=2 << 30, tt_bytes=64 << 20)
    try:
        eng.tune(pos, 4, path)
        assert np.array_equal(eng.perft_batch(pos, 4, mode=pg.MODE_PERSISTENT), want)
        assert np.array_equal(eng.perft_batch(pos, 4), want)
    finally:
        eng.close()
    text = open(path).read()
    assert text.startswith("perft_gpu_amd tuning 1") and "backend ks persistent_variant" in text
    ks = [l for l in text.splitlines() if l.startswith("backend ks ")][0]
    assert " single_pass stream " in ks or " single_pass persistent " in ks, ks
    for key in ("stream_leaf", "stream_write", "tt_leaf_blocks_per_cu", "tt_chunk_pct"):
        assert " %s " % key in ks, ks
    eng = pg.Engine(backend="ks", arena_bytes=2 << 30, tune_file=path)
    try:
        assert np.array_equal(eng.perft_batch(pos, 4, mode=pg.MODE_PERSISTENT), want)
    finally:
        eng.close()
    old = tmp_path / "old.txt"
    old.write_text("perft_gpu_amd tuning 1\ncus 256\nbackend ks persistent_variant 0 persistent_blocks_per_cu 3 "
                   "tt_blocks_per_cu 3\n")
    eng = pg.Engine(backend="ks", arena_bytes=2 << 30, tune_file=str(old))
    try:
        assert np.array_equal(eng.perft_batch(pos, 4), want)
    finally:
        eng.close()
    wrong = tmp_path / "wrong.txt"
    wrong.write_text("perft_gpu_amd tuning 1\nbackend ks persistent_variant 0 persistent_blocks_per_cu 3 "
                     "tt_blocks_per_cu 3 single_pass sideways\n")
    with pytest.raises(pg.PerftError) as ei:
        pg.Engine(backend="ks", arena_bytes=1 << 30, tune_file=str(wrong))
    assert ei.value.status == pg.PF_EINVAL
    bad = tmp_path / "bad.txt"
    bad.write_text("not a tuning file\n")
    with pytest.raises(pg.PerftError) as ei:
        pg.Engine(backend="ks", arena_bytes=1 << 30, tune_file=str(bad))
    assert ei.value.status == pg.PF_EINVAL


def test_device_resident_batches(pg, oracle, engine):
    """positions and counts in device memory (pf_perft_batch on device
    pointers): the persistent path converts on the device, every other
    combination stages through the host; all give the oracle's counts"""
    pos = pg.playouts(5150, 64, 8, 50)
    want = {d: oracle.perft_batch(pos, d, nthreads=8) for d in (1, 2, 4)}
    dpos = pg.DeviceArray.from_host(pos)
    dout = pg.DeviceArray(len(pos), np.uint64)
    hout = np.zeros(len(pos), dtype=np.uint64)
    try:
        for mode in (pg.MODE_AUTO, pg.MODE_PERSISTENT, pg.MODE_STREAM, pg.MODE_LEVELS):
            for d in (1, 2, 4):
                dout.upload(np.full(len(pos), 7, dtype=np.uint64))
                engine.perft_batch_ptr(dpos.ptr, len(pos), d, dout.ptr, mode)
                assert np.array_equal(dout.download(), want[d]), (mode, d)
        # device positions, host counts; host positions, device counts
        engine.perft_batch_ptr(dpos.ptr, len(pos), 4, hout.ctypes.data, pg.MODE_PERSISTENT)
        assert np.array_equal(hout, want[4])
        engine.perft_batch_ptr(pos.ctypes.data, len(pos), 4, dout.ptr, pg.MODE_PERSISTENT)
        assert np.array_equal(dout.download(), want[4])
        # the subset at an offset inside the device buffer
        engine.perft_batch_ptr(dpos.ptr + 48 * 10, 20, 4, dout.ptr, pg.MODE_PERSISTENT)
        assert np.array_equal(dout.download()[:20], want[4][10:30])
    finally:
        dpos.free()
        dout.free()


def test_device_resident_batch_with_table(pg, oracle):
    """with a transposition table AUTO runs the level driver: device buffers
    are staged, counts unchanged"""
    pos = pg.playouts(5151, 24, 8, 50)
    want = oracle.perft_batch(pos, 4, nthreads=8)
    eng = pg.Engine(backend="ks", arena_bytes=1 << 30, tt_bytes=64 << 20)
    dpos = pg.DeviceArray.from_host(pos)
    dout = pg.DeviceArray(len(pos), np.uint64)
    try:
        eng.perft_batch_ptr(dpos.ptr, len(pos), 4, dout.ptr, pg.MODE_AUTO)
        assert np.array_equal(dout.download(), want)
    finally:
        dpos.free()
        dout.free()
        eng.close()
