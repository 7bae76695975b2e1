// movegen_check.cpp -- TEST HARNESS ONLY (built by tests/conftest.py into
// tests/native/_build/libmovegen_check.so).  Runs the product's shared
// host/device move generator (perft_gpu_amd/csrc/movegen.h) on the CPU so
// that the CPU test suite can check it against the oracle before any GPU run.
// Not part of the product library.
#include <stdint.h>

#include "../../perft_gpu_amd/csrc/convert.h"
#include "../../perft_gpu_amd/csrc/movegen.h"
#include "../../perft_gpu_amd/csrc/tables.h"
#include "../../perft_gpu_amd/csrc/tt.h"

using namespace pf;

static SliderTables* g_tab = nullptr;

template <class SL>
static u64 rec(const Board& b, int depth, const SL& sl) {
    if (depth == 1) return count_moves(b, sl);
    PosInfo pi;
    analyse(b, sl, pi);
    u32 mv[256];
    int n = 0;
    auto emit = [&](u32 m) { mv[n++] = m; };
    generate(b, sl, pi, emit);
    u64 c = 0;
    for (int i = 0; i < n; i++) c += rec(make_move(b, mv[i]), depth - 1, sl);
    return c;
}

extern "C" {

int mg_init() {
    if (g_tab) return 0;
    g_tab = new SliderTables;
    return build_slider_tables(*g_tab);
}

uint64_t mg_perft(const pf_pos48* p, int depth, int backend) {
    if (depth <= 0) return 1;
    const Board b = to_board(*p);
    switch (backend) {
    case kPlain: return rec(b, depth, PlainBackend{g_tab->plain_b, g_tab->plain_r, g_tab->plain_btab, g_tab->plain_rtab});
    case kFancy: return rec(b, depth, FancyBackend{g_tab->fancy_b, g_tab->fancy_r, g_tab->fancy_tab});
    case kByte: return rec(b, depth, ByteBackend{g_tab->fancy_b, g_tab->fancy_r, g_tab->byte_tab, g_tab->byte_r, g_tab->byte_b});
    default: return rec(b, depth, KSBackend{});
    }
}

// generate() and count_moves() must agree on every node of a tree
int mg_gen_count_agree(const pf_pos48* p, int depth) {
    const KSBackend sl{};
    struct Walk {
        const KSBackend& sl;
        int bad = 0;
        void go(const Board& b, int d) {
            PosInfo pi;
            analyse(b, sl, pi);
            u32 mv[256];
            int n = 0;
            auto emit = [&](u32 m) { mv[n++] = m; };
            generate(b, sl, pi, emit);
            if ((u32)n != count_moves(b, sl)) bad++;
            if (d > 1)
                for (int i = 0; i < n; i++) go(make_move(b, mv[i]), d - 1);
        }
    } w{sl};
    w.go(to_board(*p), depth);
    return w.bad;
}

// attacks of one slider for table-level brute force checks
uint64_t mg_attacks(int backend, int rook, int sq, uint64_t occ) {
    switch (backend) {
    case kPlain: {
        PlainBackend s{g_tab->plain_b, g_tab->plain_r, g_tab->plain_btab, g_tab->plain_rtab};
        return rook ? s.rook(sq, occ) : s.bishop(sq, occ);
    }
    case kFancy: {
        FancyBackend s{g_tab->fancy_b, g_tab->fancy_r, g_tab->fancy_tab};
        return rook ? s.rook(sq, occ) : s.bishop(sq, occ);
    }
    case kByte: {
        ByteBackend s{g_tab->fancy_b, g_tab->fancy_r, g_tab->byte_tab, g_tab->byte_r, g_tab->byte_b};
        return rook ? s.rook(sq, occ) : s.bishop(sq, occ);
    }
    default: {
        KSBackend s{};
        return rook ? s.rook(sq, occ) : s.bishop(sq, occ);
    }
    }
}

// child positions of the product's make_move, in pf_pos48 form
int mg_children(const pf_pos48* p, pf_pos48* out) {
    const int black = (int)(p->pawns >> 63);
    const Board b = to_board(*p);
    const KSBackend sl{};
    PosInfo pi;
    analyse(b, sl, pi);
    u32 mv[256];
    int n = 0;
    auto emit = [&](u32 m) { mv[n++] = m; };
    generate(b, sl, pi, emit);
    for (int i = 0; i < n; i++) out[i] = to_pos48(make_move(b, mv[i]), !black);
    return n;
}

// full key of a position with key set `which` (0 bucket key, 1 verification key)
uint64_t mg_zkey_full(const pf_pos48* p, uint64_t seed, int which) {
    ZTables z;
    zobrist_tables(z, seed);
    return zobrist_full(z.k[which], to_board(*p), (int)(p->pawns >> 63));
}

// incremental child keys == full keys of the children, both key sets
int mg_zkey_children_check(const pf_pos48* p, uint64_t seed) {
    ZTables z;
    zobrist_tables(z, seed);
    const int black = (int)(p->pawns >> 63);
    const Board b = to_board(*p);
    const KSBackend sl{};
    PosInfo pi;
    analyse(b, sl, pi);
    u32 mv[256];
    int n = 0;
    auto emit = [&](u32 m) { mv[n++] = m; };
    generate(b, sl, pi, emit);
    int bad = 0;
    for (int w = 0; w < 2; w++) {
        const u64 k = zobrist_full(z.k[w], b, black);
        for (int i = 0; i < n; i++) {
            const Board c = make_move(b, mv[i]);
            if (zobrist_child(z.k[w], k, b, mv[i], black, c) != zobrist_full(z.k[w], c, !black)) bad++;
        }
    }
    return bad;
}

// every three-input logic helper (board.h) against the two-input expression it
// replaces, on random words: pins each truth-table immediate
int mg_bitop3_check(uint64_t seed, int n) {
    u64 x = seed ? seed : 1;
    auto nx = [&x]() { x ^= x << 13; x ^= x >> 7; x ^= x << 17; return x; };
    int bad = 0;
    for (int i = 0; i < n; i++) {
        const u64 a = nx(), b = nx(), c = nx();
        bad += or_and(a, b, c) != (a | (b & c));
        bad += or3(a, b, c) != (a | b | c);
        bad += and3(a, b, c) != (a & b & c);
        bad += and_or(a, b, c) != ((a & b) | c);
        bad += either_and(a, b, c) != ((a | b) & c);
        bad += and_andn(a, b, c) != (a & b & ~c);
        bad += andn2(a, b, c) != (a & ~b & ~c);
        bad += xor_and(a, b, c) != (a ^ (b & c));
        bad += xor_or(a, b, c) != ((a ^ b) | c);
    }
    return bad;
}
}
