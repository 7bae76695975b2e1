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


@pytest.mark.parametrize("mode", [1, 2])
def test_batch_vs_oracle(pg, oracle, engine, mode):
    arr = pg.playouts(20240601, 96, 8, 60)
    for depth in (1, 2, 3, 4):
        got = engine.perft_batch(arr, depth, mode=mode)
        want = oracle.perft_batch(arr, depth, nthreads=8)
        assert np.array_equal(got, want), (depth, mode)
