"""perft_gpu_amd -- Python host mirror of the MI355X perft engine.

Thin ctypes layer over the C ABI in include/perft_gpu_amd.h (library
perft_gpu_amd/_build/libperft_gpu_amd.so, built by perft_gpu_amd.build).
There is no CPU fallback: if the library is missing, or no GPU is usable,
every compute call raises.

Reference interface mirrored (sarona-RD/perft_gpu):
  perft_launcher(pos, depth)        perftLauncher        (launcher.h:1455-1504)
  Engine.perft_batch(positions, d)  perft_bb_gpu_simple_hash(count, ...) (perft_bb.h:1242)
  fen_to_pos(fen)                   Utils::readFENString + board088ToHexBB (util.cpp:189, :92)
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libperft_gpu_amd.so")
CLI_PATH = os.path.join(HERE, "_build", "perft_gpu")

PF_OK, PF_EINVAL, PF_ENOMEM, PF_EDEVICE, PF_EBUSY, PF_EAGAIN, PF_EUNAVAIL, PF_EIO = 0, -1, -2, -3, -4, -5, -6, -7
BACKENDS = {"ks": 0, "plain": 1, "fancy": 2, "byte": 3}
MODE_AUTO, MODE_LEVELS, MODE_PERSISTENT, MODE_STREAM = 0, 1, 2, 3
TUNE_ALL, TUNE_SINGLE_PASS, TUNE_TABLE = 0, 1, 2
START_FEN = "rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1"

# structured dtype with the exact pf_pos48 layout, for zero-copy batches
POS_DTYPE = np.dtype([("white_pieces", "<u8"), ("pawns", "<u8"), ("knights", "<u8"),
                      ("bishop_queens", "<u8"), ("rook_queens", "<u8"), ("kings", "<u8")])


class PerftError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__("%s: status %d (%s)" % (what or "perft_gpu_amd", status, _strerror(status)))


class Pos48(ctypes.Structure):
    _fields_ = [("white_pieces", ctypes.c_uint64), ("pawns", ctypes.c_uint64), ("knights", ctypes.c_uint64),
                ("bishop_queens", ctypes.c_uint64), ("rook_queens", ctypes.c_uint64), ("kings", ctypes.c_uint64)]

    def as_tuple(self):
        return (self.white_pieces, self.pawns, self.knights, self.bishop_queens, self.rook_queens, self.kings)


class Config(ctypes.Structure):
    _fields_ = [("backend", ctypes.c_int), ("device", ctypes.c_int), ("arena_bytes", ctypes.c_uint64),
                ("tt_bytes", ctypes.c_uint64), ("zobrist_seed", ctypes.c_uint64), ("tune_file", ctypes.c_char_p)]


class Result(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("nodes", ctypes.c_uint64), ("status", ctypes.c_int32),
                ("depth", ctypes.c_int32), ("submit_s", ctypes.c_double), ("done_s", ctypes.c_double)]


class Stats(ctypes.Structure):
    _fields_ = [("kernel_launches", ctypes.c_uint64), ("levels_run", ctypes.c_uint64), ("chunks", ctypes.c_uint64),
                ("peak_arena_bytes", ctypes.c_uint64), ("tt_probes", ctypes.c_uint64), ("tt_hits", ctypes.c_uint64),
                ("tt_stores", ctypes.c_uint64), ("gpu_ms", ctypes.c_double), ("last_expand_ms", ctypes.c_double),
                ("last_expand_launches", ctypes.c_uint64), ("dup_nodes", ctypes.c_uint64),
                ("frontier_nodes", ctypes.c_uint64), ("made_nodes", ctypes.c_uint64),
                ("graph_passes", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

# every symbol the C ABI declares (checked by the CPU test suite)
EXPORTS = ["pf_init", "pf_destroy", "pf_strerror", "pf_fen_to_pos", "pf_pos_to_fen", "pf_perft",
           "pf_perft_batch", "pf_perft_batch_mode", "pf_perft_batch_noalloc", "pf_divide", "pf_move_to_uci", "pf_apply_uci",
           "pf_set_backend", "pf_set_dedup", "pf_set_level_graphs", "pf_set_exact_ep", "pf_set_tt_regions", "pf_tune", "pf_tune_mode", "pf_records", "pf_tt_save", "pf_tt_load", "pf_get_stats", "pf_reset_stats",
           "pf_sync", "pf_probe", "pf_playouts", "pf_count_moves_host", "pf_serve_start", "pf_submit", "pf_poll", "pf_serve_stop", "pf_now"]


def load_library(path=LIB_PATH):
    """Load the engine library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("perft_gpu_amd library not built (%s); run `python -m perft_gpu_amd.build`" % path)
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    u64p = P(ctypes.c_uint64)
    sig = {
        "pf_init": (ctypes.c_int, [P(Config), P(ctypes.c_void_p)]),
        "pf_destroy": (None, [ctypes.c_void_p]),
        "pf_strerror": (ctypes.c_char_p, [ctypes.c_int]),
        "pf_fen_to_pos": (ctypes.c_int, [ctypes.c_char_p, P(Pos48)]),
        "pf_pos_to_fen": (ctypes.c_int, [P(Pos48), ctypes.c_char_p, ctypes.c_int]),
        "pf_perft": (ctypes.c_int, [ctypes.c_void_p, P(Pos48), ctypes.c_int, u64p]),
        "pf_perft_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, u64p]),
        "pf_perft_batch_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, u64p]),
        "pf_perft_batch_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_void_p]),
        "pf_divide": (ctypes.c_int, [ctypes.c_void_p, P(Pos48), ctypes.c_int, P(ctypes.c_uint16), u64p,
                                     P(ctypes.c_int)]),
        "pf_move_to_uci": (ctypes.c_int, [ctypes.c_uint16, ctypes.c_char_p]),
        "pf_apply_uci": (ctypes.c_int, [P(Pos48), ctypes.c_char_p]),
        "pf_set_backend": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "pf_get_stats": (ctypes.c_int, [ctypes.c_void_p, P(Stats)]),
        "pf_reset_stats": (ctypes.c_int, [ctypes.c_void_p]),
        "pf_sync": (ctypes.c_int, [ctypes.c_void_p]),
        "pf_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                                    ctypes.POINTER(ctypes.c_double)]),
        "pf_playouts": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_void_p]),
        "pf_count_moves_host": (ctypes.c_int, [P(Pos48)]),
        "pf_set_dedup": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "pf_set_level_graphs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "pf_set_exact_ep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "pf_set_tt_regions": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "pf_tune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p]),
        "pf_tune_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_char_p]),
        "pf_records": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, u64p, u64p]),
        "pf_tt_save": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
        "pf_tt_load": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
        "pf_submit": (ctypes.c_int, [ctypes.c_void_p, P(Pos48), ctypes.c_int, ctypes.c_uint64]),
        "pf_poll": (ctypes.c_int, [ctypes.c_void_p, P(Result), ctypes.c_int]),
        "pf_serve_start": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64]),
        "pf_serve_stop": (ctypes.c_int, [ctypes.c_void_p]),
        "pf_now": (ctypes.c_double, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _strerror(status):
    try:
        return load_library().pf_strerror(status).decode()
    except Exception:
        return "?"


def _check(status, what):
    if status != PF_OK:
        raise PerftError(status, what)


def fen_to_pos(fen):
    p = Pos48()
    _check(load_library().pf_fen_to_pos(fen.encode(), ctypes.byref(p)), "fen_to_pos(%r)" % fen)
    return p


def pos_to_fen(pos):
    buf = ctypes.create_string_buffer(128)
    _check(load_library().pf_pos_to_fen(ctypes.byref(pos), buf, 128), "pos_to_fen")
    return buf.value.decode()


def move_to_uci(move):
    buf = ctypes.create_string_buffer(8)
    load_library().pf_move_to_uci(move, buf)
    return buf.value.decode()


def apply_uci(pos, uci):
    p = Pos48(*pos.as_tuple())
    _check(load_library().pf_apply_uci(ctypes.byref(p), uci.encode()), "apply_uci(%r)" % uci)
    return p


def positions_array(positions):
    """list of Pos48 / structured array -> contiguous POS_DTYPE array"""
    if isinstance(positions, np.ndarray) and positions.dtype == POS_DTYPE:
        return np.ascontiguousarray(positions)
    arr = np.zeros(len(positions), dtype=POS_DTYPE)
    for i, p in enumerate(positions):
        arr[i] = p.as_tuple() if hasattr(p, "as_tuple") else tuple(int(x) for x in p)
    return arr


def playouts(seed, n, min_plies, max_plies, min_moves=0, no_check=False):
    arr = np.zeros(n, dtype=POS_DTYPE)
    _check(load_library().pf_playouts(seed, n, min_plies, max_plies, min_moves, int(bool(no_check)),
                                      arr.ctypes.data_as(ctypes.c_void_p)), "playouts")
    return arr


_hip_rt = None


def _hip():
    """The HIP runtime the engine library itself links (same loaded copy)."""
    global _hip_rt
    if _hip_rt is None:
        load_library()
        rt = ctypes.CDLL("libamdhip64.so.7")
        for name, args in (("hipSetDevice", [ctypes.c_int]),
                           ("hipMalloc", [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]),
                           ("hipFree", [ctypes.c_void_p]),
                           ("hipMemcpy", [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int])):
            fn = getattr(rt, name)
            fn.restype = ctypes.c_int
            fn.argtypes = args
        _hip_rt = rt
    return _hip_rt


def _hip_check(err, what):
    if err != 0:
        raise RuntimeError("%s: HIP error %d" % (what, err))


class DeviceArray:
    """n elements of a numpy dtype in device memory (hipMalloc), for batches
    that stay resident in HBM between calls; `ptr` is the device address."""

    def __init__(self, n, dtype, device=0):
        self.dtype, self.n, self.device = np.dtype(dtype), int(n), device
        self.nbytes = self.n * self.dtype.itemsize
        hip, p = _hip(), ctypes.c_void_p()
        _hip_check(hip.hipSetDevice(device), "hipSetDevice")
        _hip_check(hip.hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc(%d)" % self.nbytes)
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr, device=0):
        arr = np.ascontiguousarray(arr)
        d = cls(len(arr), arr.dtype, device)
        d.upload(arr)
        return d

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes
        _hip_check(_hip().hipMemcpy(self.ptr, arr.ctypes.data, self.nbytes, 1), "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        _hip_check(_hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        if getattr(self, "ptr", None):
            _hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Engine:
    """One device context (pf_ctx).  Mirrors the reference's GPU perft entry
    points; every compute call runs the HIP kernels."""

    def __init__(self, backend="ks", device=0, arena_bytes=0, tt_bytes=0, zobrist_seed=0, tune_file=None):
        lib = load_library()
        self._lib = lib
        cfg = Config(BACKENDS[backend] if isinstance(backend, str) else int(backend), device, arena_bytes,
                     tt_bytes, zobrist_seed, tune_file.encode() if tune_file else None)
        ctx = ctypes.c_void_p()
        _check(lib.pf_init(ctypes.byref(cfg), ctypes.byref(ctx)), "pf_init")
        self._ctx = ctx

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.pf_destroy(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_backend(self, backend):
        _check(self._lib.pf_set_backend(self._ctx, BACKENDS[backend] if isinstance(backend, str) else int(backend)),
               "set_backend")

    def tune(self, positions, depth, path=None, what=TUNE_ALL):
        """Autotune the current backend's launch parameters on a sample
        (pf_tune_mode: TUNE_ALL, TUNE_SINGLE_PASS or TUNE_TABLE); optionally
        write the tuning file."""
        arr = positions_array(positions)
        _check(self._lib.pf_tune_mode(self._ctx, arr.ctypes.data, len(arr), depth, what,
                                      path.encode() if path else None), "tune")

    def records(self, path, depth, batch=0):
        """Records mode (pf_records): returns (weighted_total, records_done)."""
        total, done = ctypes.c_uint64(), ctypes.c_uint64()
        _check(self._lib.pf_records(self._ctx, path.encode(), depth, batch, ctypes.byref(total), ctypes.byref(done)),
               "records(%r)" % path)
        return total.value, done.value

    def tt_save(self, path):
        """Write the transposition table to `path` (pf_tt_save)."""
        _check(self._lib.pf_tt_save(self._ctx, path.encode()), "tt_save(%r)" % path)

    def tt_load(self, path):
        """Replace the transposition table with a saved one (pf_tt_load)."""
        _check(self._lib.pf_tt_load(self._ctx, path.encode()), "tt_load(%r)" % path)

    def set_level_graphs(self, on):
        """pf_set_level_graphs: table runs use captured-graph passes (default on)."""
        _check(self._lib.pf_set_level_graphs(self._ctx, 1 if on else 0), "set_level_graphs")

    def set_exact_ep(self, on):
        """pf_set_exact_ep: children flag en passant only when capturable (default)."""
        _check(self._lib.pf_set_exact_ep(self._ctx, 1 if on else 0), "set_exact_ep")

    def set_tt_regions(self, log2_regions):
        """pf_set_tt_regions: 0 one table for every depth, k: 2^k per-depth regions."""
        _check(self._lib.pf_set_tt_regions(self._ctx, log2_regions), "set_tt_regions")

    def set_dedup(self, on):
        _check(self._lib.pf_set_dedup(self._ctx, 1 if on else 0), "set_dedup")

    def perft(self, pos, depth):
        out = ctypes.c_uint64()
        _check(self._lib.pf_perft(self._ctx, ctypes.byref(pos), depth, ctypes.byref(out)), "perft")
        return out.value

    def perft_batch(self, positions, depth, mode=MODE_AUTO):
        arr = positions_array(positions)
        out = np.zeros(len(arr), dtype=np.uint64)
        _check(self._lib.pf_perft_batch_mode(self._ctx, arr.ctypes.data_as(ctypes.c_void_p), len(arr), depth, mode,
                                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))), "perft_batch")
        return out

    def perft_batch_ptr(self, pos_ptr, n, depth, out_ptr, mode=MODE_AUTO):
        """pf_perft_batch_mode on raw addresses: `pos_ptr` -> n pf_pos48 records,
        `out_ptr` -> n uint64 counts, each host or device memory of this
        context's device (e.g. a torch tensor's data_ptr())."""
        _check(self._lib.pf_perft_batch_mode(self._ctx, ctypes.c_void_p(pos_ptr), n, depth, mode,
                                             ctypes.cast(ctypes.c_void_p(out_ptr), ctypes.POINTER(ctypes.c_uint64))),
               "perft_batch_ptr")

    def perft_batch_noalloc_ptr(self, pos_ptr, n, depth, out_ptr):
        """pf_perft_batch_noalloc: the stream-ordered batch on device buffers
        (`pos_ptr`, `out_ptr` device addresses of this context's device),
        allocating nothing -- its levels live in the library's fixed 512 MiB
        device array; PerftError PF_ENOMEM when they do not fit."""
        _check(self._lib.pf_perft_batch_noalloc(self._ctx, ctypes.c_void_p(pos_ptr), n, depth,
                                                ctypes.c_void_p(out_ptr)), "perft_batch_noalloc_ptr")

    def divide(self, pos, depth):
        mv = (ctypes.c_uint16 * 256)()
        cnt = (ctypes.c_uint64 * 256)()
        n = ctypes.c_int()
        _check(self._lib.pf_divide(self._ctx, ctypes.byref(pos), depth, mv, cnt, ctypes.byref(n)), "divide")
        return [(move_to_uci(mv[i]), cnt[i]) for i in range(n.value)]

    def stats(self):
        s = Stats()
        _check(self._lib.pf_get_stats(self._ctx, ctypes.byref(s)), "stats")
        return s.as_dict()

    def sync(self):
        _check(self._lib.pf_sync(self._ctx), "sync")

    def probe(self, kind, nbytes=0, iters=0):
        """measured ceiling: 0 int32 lane-ops/s, 1 copy bytes/s, 2 random 16-B RMW/s"""
        out = ctypes.c_double()
        _check(self._lib.pf_probe(self._ctx, kind, nbytes, iters, ctypes.byref(out)), "probe")
        return out.value

    # ---- serving (continuous batching) ----
    def serve_start(self, max_jobs=64, job_bytes=0):
        _check(self._lib.pf_serve_start(self._ctx, max_jobs, job_bytes), "serve_start")

    def submit(self, pos, depth, tag):
        p = pos if isinstance(pos, Pos48) else Pos48(*[int(x) for x in pos])
        _check(self._lib.pf_submit(self._ctx, ctypes.byref(p), depth, tag), "submit")

    def poll(self, max_results=4096):
        buf = (Result * max_results)()
        n = self._lib.pf_poll(self._ctx, buf, max_results)
        if n < 0:
            raise PerftError(n, "poll")
        return [{"tag": buf[i].tag, "nodes": buf[i].nodes, "status": buf[i].status, "depth": buf[i].depth,
                 "submit_s": buf[i].submit_s, "done_s": buf[i].done_s} for i in range(n)]

    def serve_stop(self):
        _check(self._lib.pf_serve_stop(self._ctx), "serve_stop")

    def now(self):
        return self._lib.pf_now()

    def reset_stats(self):
        _check(self._lib.pf_reset_stats(self._ctx), "reset_stats")


def perft_launcher(pos, depth, engine=None):
    """perftLauncher(pos, depth, launchDepth) (launcher.h:1455): returns the
    count instead of printing it; launch depth is chosen by the engine."""
    own = engine is None
    eng = engine or Engine()
    try:
        return eng.perft(pos, depth)
    finally:
        if own:
            eng.close()
