"""ctypes wrapper of the oracle (oracle/_build/liboracle.so) -- test
infrastructure only; the product never imports this module."""
import ctypes
import os
import subprocess
import threading

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
LIB_OPS = os.path.join(ORACLE_DIR, "_build", "liboracle_ops.so")

KS, PLAIN, FANCY, BYTE = 0, 1, 2, 3


class OPos(ctypes.Structure):
    _fields_ = [("white", ctypes.c_uint64), ("pawns", ctypes.c_uint64), ("knights", ctypes.c_uint64),
                ("bishop_queens", ctypes.c_uint64), ("rook_queens", ctypes.c_uint64), ("kings", ctypes.c_uint64)]

    def as_tuple(self):
        return (self.white, self.pawns, self.knights, self.bishop_queens, self.rook_queens, self.kings)


def ensure_built():
    if not (os.path.exists(LIB) and os.path.exists(LIB_OPS)):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def _load(path):
    ensure_built()
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    u64 = ctypes.c_uint64
    lib.orc_init.restype = ctypes.c_int
    lib.orc_set_backend.argtypes = [ctypes.c_int]
    lib.orc_fen_to_pos.argtypes = [ctypes.c_char_p, P(OPos)]
    lib.orc_pos_to_fen.argtypes = [P(OPos), ctypes.c_char_p, ctypes.c_int]
    lib.orc_count_moves.argtypes = [P(OPos)]
    lib.orc_count_moves.restype = ctypes.c_uint32
    lib.orc_generate_boards.argtypes = [P(OPos), P(OPos)]
    lib.orc_generate_boards.restype = ctypes.c_uint32
    lib.orc_generate_moves.argtypes = [P(OPos), P(ctypes.c_uint16)]
    lib.orc_generate_moves.restype = ctypes.c_uint32
    lib.orc_make_move.argtypes = [P(OPos), ctypes.c_uint16, ctypes.c_int]
    lib.orc_in_check.argtypes = [P(OPos)]
    lib.orc_perft.argtypes = [P(OPos), ctypes.c_int]
    lib.orc_perft.restype = u64
    lib.orc_perft_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(u64)]
    lib.orc_divide.argtypes = [P(OPos), ctypes.c_int, P(ctypes.c_uint16), P(u64)]
    lib.orc_rook_attacks.argtypes = [ctypes.c_int, ctypes.c_int, u64]
    lib.orc_rook_attacks.restype = u64
    lib.orc_bishop_attacks.argtypes = [ctypes.c_int, ctypes.c_int, u64]
    lib.orc_bishop_attacks.restype = u64
    lib.orc_rook_mask.argtypes = [ctypes.c_int]
    lib.orc_rook_mask.restype = u64
    lib.orc_bishop_mask.argtypes = [ctypes.c_int]
    lib.orc_bishop_mask.restype = u64
    lib.orc_xorshift64.argtypes = [P(u64)]
    lib.orc_xorshift64.restype = u64
    lib.orc_zobrist_table.argtypes = [u64, P(u64)]
    lib.orc_zobrist_key.argtypes = [P(OPos), P(u64)]
    lib.orc_zobrist_key.restype = u64
    lib.orc_playouts.argtypes = [u64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_void_p]
    lib.orc_opcount_perft.argtypes = [P(OPos), ctypes.c_int, P(u64), P(u64)]
    lib.orc_naive_perft.argtypes = [P(OPos), ctypes.c_int]
    lib.orc_naive_perft.restype = u64
    rc = lib.orc_init()
    if rc != 0:
        raise RuntimeError("oracle init failed: %d" % rc)
    return lib


class Oracle:
    _lib = None
    _lib_ops = None
    _lock = threading.Lock()  # orc_init is not re-entrant

    def __init__(self):
        with Oracle._lock:
            if Oracle._lib is None:
                Oracle._lib = _load(LIB)
        self.lib = Oracle._lib

    @property
    def ops_lib(self):
        with Oracle._lock:
            if Oracle._lib_ops is None:
                Oracle._lib_ops = _load(LIB_OPS)
        return Oracle._lib_ops

    def set_backend(self, b):
        return self.lib.orc_set_backend(b)

    def pos(self, fen):
        p = OPos()
        if self.lib.orc_fen_to_pos(fen.encode(), ctypes.byref(p)) != 0:
            raise ValueError("bad FEN %r" % fen)
        return p

    @staticmethod
    def from48(p48):
        """accept a Pos48 / structured-array row / 6-tuple"""
        if hasattr(p48, "as_tuple"):
            return OPos(*p48.as_tuple())
        return OPos(*[int(x) for x in tuple(p48)])

    def fen(self, p):
        buf = ctypes.create_string_buffer(128)
        self.lib.orc_pos_to_fen(ctypes.byref(p), buf, 128)
        return buf.value.decode()

    def perft(self, p, depth):
        return self.lib.orc_perft(ctypes.byref(p), depth)

    def perft_fen(self, fen, depth):
        return self.perft(self.pos(fen), depth)

    def perft_pos48(self, p48, depth):
        return self.perft(self.from48(p48), depth)

    def perft_batch(self, positions, depth, nthreads=1):
        arr = np.ascontiguousarray(positions)
        out = np.zeros(len(arr), dtype=np.uint64)
        self.lib.orc_perft_batch(arr.ctypes.data_as(ctypes.c_void_p), len(arr), depth, nthreads,
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        return out

    def count(self, p):
        return self.lib.orc_count_moves(ctypes.byref(p))

    def moves(self, p):
        buf = (ctypes.c_uint16 * 256)()
        n = self.lib.orc_generate_moves(ctypes.byref(p), buf)
        return list(buf[:n])

    def children(self, p):
        buf = (OPos * 256)()
        n = self.lib.orc_generate_boards(ctypes.byref(p), buf)
        return [buf[i] for i in range(n)]

    def make(self, p, m, exact_ep=0):
        c = OPos(*p.as_tuple())
        self.lib.orc_make_move(ctypes.byref(c), m, exact_ep)
        return c

    def in_check(self, p):
        return bool(self.lib.orc_in_check(ctypes.byref(p)))

    def divide(self, p, depth):
        mv = (ctypes.c_uint16 * 256)()
        cnt = (ctypes.c_uint64 * 256)()
        n = self.lib.orc_divide(ctypes.byref(p), depth, mv, cnt)
        return [(mv[i], cnt[i]) for i in range(n)]

    def naive_perft(self, p, depth):
        return self.lib.orc_naive_perft(ctypes.byref(p), depth)

    def rook_attacks(self, backend, sq, occ):
        return self.lib.orc_rook_attacks(backend, sq, occ)

    def bishop_attacks(self, backend, sq, occ):
        return self.lib.orc_bishop_attacks(backend, sq, occ)

    def zobrist_table(self, seed):
        t = (ctypes.c_uint64 * 781)()
        self.lib.orc_zobrist_table(seed, t)
        return t

    def zobrist_key(self, p, table):
        return self.lib.orc_zobrist_key(ctypes.byref(p), table)

    def playouts(self, seed, n, lo, hi, min_moves=0, no_check=False):
        arr = np.zeros(n, dtype=[("white", "<u8"), ("pawns", "<u8"), ("knights", "<u8"),
                                 ("bishop_queens", "<u8"), ("rook_queens", "<u8"), ("kings", "<u8")])
        got = self.lib.orc_playouts(seed, n, lo, hi, min_moves, int(bool(no_check)),
                                    arr.ctypes.data_as(ctypes.c_void_p))
        return arr[:got]

    def opcount(self, p, depth):
        ops = ctypes.c_uint64()
        leaves = ctypes.c_uint64()
        self.ops_lib.orc_opcount_perft(ctypes.byref(p), depth, ctypes.byref(ops), ctypes.byref(leaves))
        return ops.value, leaves.value
