"""Records mode (pf_records; processPerftRecords, launcher.h:1578-1695): every
"<FEN> <occurrences>" line gets "<line> <perft> <perft*occurrences>" in
<file>.op, in input order, and an interrupted run resumes where it stopped."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu


def _records(pg, oracle, tmp_path, n=20, depth=4):
    pos = pg.playouts(2468, n, 6, 50)
    lines, want = [], 0
    for i, r in enumerate(pos):
        fen = pg.pos_to_fen(pg.Pos48(*[int(x) for x in r]))
        occ = 1 + (i * 7) % 13
        lines.append("%s %d" % (fen, occ))
        want += int(oracle.perft_batch(pos[i:i + 1], depth, nthreads=1)[0]) * occ
    path = tmp_path / "records.txt"
    path.write_text("\n".join(lines[:10]) + "\n\n" + "\n".join(lines[10:]) + "\n")  # a blank line is skipped
    return str(path), lines, want


def test_records_weighted_total_and_format(pg, oracle, tmp_path):
    path, lines, want = _records(pg, oracle, tmp_path)
    eng = pg.Engine(backend="ks", arena_bytes=2 << 30, tt_bytes=64 << 20)
    try:
        total, done = eng.records(path, 4, batch=7)
    finally:
        eng.close()
    assert (total, done) == (want, 20)
    out = open(path + ".op").read().splitlines()
    assert len(out) == 20
    for line, o in zip(lines, out):
        assert o.startswith(line + " ")
        perft, weighted = map(int, o[len(line) + 1:].split())
        assert weighted == perft * int(line.split()[-1])


def test_records_resume_after_interruption(pg, oracle, tmp_path):
    path, lines, want = _records(pg, oracle, tmp_path)
    eng = pg.Engine(backend="fancy", arena_bytes=2 << 30)
    try:
        eng.records(path, 4)
        full = open(path + ".op").read()
        kept = full.splitlines(keepends=True)[:5]
        with open(path + ".op", "w") as f:  # five finished records and a torn sixth
            f.write("".join(kept) + lines[5][:17])
        total, done = eng.records(path, 4, batch=4)
        assert (total, done) == (want, 15)
        assert open(path + ".op").read() == full
        assert eng.records(path, 4) == (want, 0)  # nothing left to do
        with open(path + ".op", "w") as f:  # an output file of another input
            f.write("8/8/8/8/8/8/8/K6k w - - 0 1 3 3 9\n")
        with pytest.raises(pg.PerftError) as ei:
            eng.records(path, 4)
        assert ei.value.status == pg.PF_EINVAL
    finally:
        eng.close()


def test_records_cli(pg, oracle, tmp_path):
    path, _, want = _records(pg, oracle, tmp_path, n=12, depth=3)
    r = subprocess.run([pg.CLI_PATH, "records", path, "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().endswith("weighted perft 3: %d" % want), r.stdout
