"""Saved and restored transposition tables (pf_tt_save / pf_tt_load; the
reference's disk hash diskTTProbe / diskTTStore, launcher.h:700-789, which
keeps deep subtree counts across runs).  A restored table -- same size,
folded into a smaller one, replicated into a larger one -- must change no
count (the oracle's), and a warm run must hit it."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIB = 1 << 20


@pytest.fixture(scope="module")
def case(oracle, pg):
    pos = pg.playouts(90210, 12, 10, 40, min_moves=20)
    return pos, oracle.perft_batch(pos, 5, nthreads=8)


@pytest.fixture(scope="module")
def kids(oracle, pg, case, engine):
    """every child of the roots, made by the engine's own move maker: a cold
    perft 5 of the roots stores each of them at remaining depth 4"""
    pos, _ = case
    out = []
    for row in pos:
        p = pg.Pos48(*[int(x) for x in row])
        out += [pg.apply_uci(p, m) for m, _ in engine.divide(p, 1)]
    arr = pg.positions_array(out)
    return arr, oracle.perft_batch(arr, 4, nthreads=8)


def _engine(pg, tt_mib, seed=0):
    return pg.Engine(backend="ks", arena_bytes=2 << 30, tt_bytes=tt_mib * MIB, zobrist_seed=seed)


@pytest.fixture(scope="module")
def saved(pg, case, tmp_path_factory):
    """a 64 MiB table after one cold run, saved; the cold run's stats"""
    pos, want = case
    path = str(tmp_path_factory.mktemp("tt") / "table.pftt")
    eng = _engine(pg, 64)
    try:
        assert np.array_equal(eng.perft_batch(pos, 5), want)
        cold = eng.stats()
        eng.tt_save(path)
    finally:
        eng.close()
    return path, cold


def test_saved_file_layout(saved):
    path, _ = saved
    assert os.path.getsize(path) == 64 + 64 * MIB  # header + 2^21 buckets x 32 B
    with open(path, "rb") as f:
        head = f.read(64)
    assert head[:8] == b"PFTTABLE"
    assert int.from_bytes(head[16:24], "little") == 64 * MIB // 32


@pytest.mark.parametrize("tt_mib", [64, 16, 256])
def test_restored_table_keeps_counts_and_hits(pg, case, kids, saved, tt_mib):
    """same size: exact restore; 16 MiB: folded; 256 MiB: replicated.  The
    roots are answered from their own entries, their children (a new batch)
    from the entries stored one ply down"""
    pos, want = case
    kpos, kwant = kids
    path, cold = saved
    eng = _engine(pg, tt_mib)
    try:
        eng.tt_load(path)
        eng.reset_stats()
        assert np.array_equal(eng.perft_batch(kpos, 4), kwant), tt_mib
        kst = eng.stats()
        eng.tt_load(path)
        eng.reset_stats()
        assert np.array_equal(eng.perft_batch(pos, 5), want), tt_mib
        warm = eng.stats()
    finally:
        eng.close()
    assert warm["tt_hits"] == len(pos) and warm["frontier_nodes"] == 0, (tt_mib, warm)
    assert cold["tt_hits"] == 0 and cold["frontier_nodes"] > 0
    # children: restored at this size or larger, and kept when folding into a
    # quarter of the buckets (each bucket keeps its two deepest entries, and
    # a few hundred depth-4 entries rarely meet in 2^19 buckets); the margin
    # covers a cold-run store that displaced another in a shared bucket
    assert kst["tt_hits"] >= 0.95 * len(kpos), (tt_mib, kst, len(kpos))


def test_load_replaces_contents(pg, case, saved, tmp_path):
    """loading an empty table's file over a restored one empties it again:
    the next run behaves exactly as the cold run did"""
    pos, want = case
    path, cold = saved
    empty = str(tmp_path / "empty.pftt")
    eng = _engine(pg, 64)
    try:
        eng.tt_save(empty)
        eng.tt_load(path)
        eng.tt_load(empty)
        eng.reset_stats()
        assert np.array_equal(eng.perft_batch(pos, 5), want)
        st = eng.stats()
    finally:
        eng.close()
    assert (st["tt_hits"], st["frontier_nodes"]) == (cold["tt_hits"], cold["frontier_nodes"])


def test_rejects_foreign_and_damaged_files(pg, saved, tmp_path):
    path, _ = saved
    other = _engine(pg, 64, seed=12345)
    try:
        with pytest.raises(pg.PerftError) as ei:  # other Zobrist keys
            other.tt_load(path)
        assert ei.value.status == pg.PF_EINVAL
    finally:
        other.close()
    trunc = tmp_path / "trunc.pftt"
    with open(path, "rb") as f:
        trunc.write_bytes(f.read(64 + 4096))
    bad = tmp_path / "bad.pftt"
    bad.write_bytes(b"NOTATABL" + bytes(56 + 32 * 1024))
    eng = _engine(pg, 64)
    try:
        for p, status in ((str(trunc), pg.PF_EINVAL), (str(bad), pg.PF_EINVAL),
                          (str(tmp_path / "missing.pftt"), pg.PF_EIO)):
            with pytest.raises(pg.PerftError) as ei:
                eng.tt_load(p)
            assert ei.value.status == status, p
        with pytest.raises(pg.PerftError) as ei:
            eng.tt_save(str(tmp_path / "no" / "such" / "dir.pftt"))
        assert ei.value.status == pg.PF_EIO
    finally:
        eng.close()
    plain = pg.Engine(backend="ks", arena_bytes=1 << 30)
    try:
        with pytest.raises(pg.PerftError) as ei:  # no table in this context
            plain.tt_save(str(tmp_path / "x.pftt"))
        assert ei.value.status == pg.PF_EINVAL
    finally:
        plain.close()


def test_cli_table_file_round_trip(pg, tmp_path):
    """perft_gpu ... <tt_mib> <file>: the second run restores the first's table"""
    path = str(tmp_path / "cli.pftt")
    cmd = [pg.CLI_PATH, "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 0 1", "4", "ks", "32", path]
    runs = [subprocess.run(cmd, capture_output=True, text=True, timeout=120) for _ in range(2)]
    for r in runs:
        assert r.returncode == 0, r.stderr
        counts = [int(m) for m in re.findall(r"GPU Perft \d+: (\d+),", r.stdout)]
        assert counts == [48, 2039, 97862, 4085603]
    assert os.path.getsize(path) == 64 + 32 * MIB
