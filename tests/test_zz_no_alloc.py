"""pf_perft_batch_noalloc: the stream-ordered batch with its control block,
roots and levels in the library's fixed device array, checked against the
existing device-buffer entry point (pf_perft_batch_mode, PF_MODE_STREAM) on
the same inputs -- two calls in a row, the second smaller -- and against the
oracle's golden counts."""
import ctypes
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "workloads.json")


def _noalloc(pg):
    fn = pg.load_library().pf_perft_batch_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    return fn


def test_noalloc_equals_stream_entry_point(pg):
    from perft_gpu_amd import workloads as W
    with open(GOLDEN) as f:
        gold = json.load(f)["c2"]["perft"]
    pos = W.positions(W.C2)
    fn = _noalloc(pg)
    eng = pg.Engine(backend="ks", arena_bytes=2 << 30)
    try:
        # the full C2 batch at perft 5, then a smaller one at perft 4 (the array
        # keeps the first call's contents: the second must not see them)
        for n, depth in ((256, 5), (40, 4)):
            print("noalloc case: %d positions, perft %d" % (n, depth), flush=True)
            dpos = pg.DeviceArray.from_host(pos[:n])
            dref = pg.DeviceArray(n, np.uint64)
            dout = pg.DeviceArray(n, np.uint64)
            eng.perft_batch_ptr(dpos.ptr, n, depth, dref.ptr, pg.MODE_STREAM)
            assert fn(eng._ctx, dpos.ptr, n, depth, dout.ptr) == pg.PF_OK
            ref, got = dref.download(), dout.download()
            assert got.tolist() == ref.tolist()
            assert got.tolist() == gold[str(depth)][:n]
            for a in (dpos, dref, dout):
                a.free()
    finally:
        eng.close()


def test_noalloc_rejects_what_does_not_fit(pg):
    from perft_gpu_amd import workloads as W
    pos = W.positions(W.C2)
    fn = _noalloc(pg)
    eng = pg.Engine(backend="ks", arena_bytes=1 << 30)
    try:
        dpos = pg.DeviceArray.from_host(pos[:64])
        dout = pg.DeviceArray(64, np.uint64)
        # perft 6 of 64 positions: about 2e9 frontier bytes, beyond the array
        print("noalloc case: 64 positions, perft 6 (does not fit)", flush=True)
        assert fn(eng._ctx, dpos.ptr, 64, 6, dout.ptr) == pg.PF_ENOMEM
        # host buffers are refused; the next call still runs
        host = np.zeros(4, dtype=np.uint64)
        assert fn(eng._ctx, dpos.ptr, 4, 3, host.ctypes.data) == pg.PF_EINVAL
        print("noalloc case: 64 positions, perft 3 after the refusal", flush=True)
        assert fn(eng._ctx, dpos.ptr, 64, 3, dout.ptr) == pg.PF_OK
        assert dout.download().tolist() == eng.perft_batch(pos[:64], 3).tolist()
        dpos.free()
        dout.free()
    finally:
        eng.close()
