// tables.cpp -- host construction of the slider tables (product side).
//
// MoveGeneratorBitboard.h:1024-1080 and Magics.cpp:42-188 build the same
// kinds of tables: masks with the board edge removed, a magic search for the
// plain tables, and a verified fill of the shared fixed-shift table plus the
// byte-lookup dedupe.  Attack sets come from the Kogge-Stone fills in
// board.h, so every backend is checked against Kogge-Stone while it is built:
// a destructive collision makes build_slider_tables() fail.
#include <string.h>

#include "tables.h"

namespace pf {

static u64 subset(u64 mask, u32 i) {  // i-th subset of mask, bit k of i -> k-th lowest bit
    u64 r = 0;
    while (i) {
        const u64 b = mask & (0 - mask);
        mask ^= b;
        if (i & 1) r |= b;
        i >>= 1;
    }
    return r;
}

static u64 xs(u64& s) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    return s;
}

u64 rook_mask(int sq) {
    const u64 t = bit(sq);
    u64 m = rook_fill(t, ~0ULL);
    if (!(t & kRank1)) m &= ~kRank1;
    if (!(t & kRank8)) m &= ~kRank8;
    if (!(t & kFileA)) m &= ~kFileA;
    if (!(t & kFileH)) m &= ~kFileH;
    return m;
}

u64 bishop_mask(int sq) { return bishop_fill(bit(sq), ~0ULL) & 0x007E7E7E7E7E7E00ULL; }

static bool find_plain(int sq, bool rook, u64& seed, MagicEntry& e, u64* win) {
    const int bits = rook ? 12 : 9;
    const u64 mask = rook ? rook_mask(sq) : bishop_mask(sq);
    const int n = 1 << popc(mask);
    static u64 occ[4096], att[4096];
    for (int i = 0; i < n; i++) {
        occ[i] = subset(mask, (u32)i);
        att[i] = rook ? rook_fill(bit(sq), ~occ[i]) : bishop_fill(bit(sq), ~occ[i]);
    }
    for (int tries = 0; tries < 10000000; tries++) {
        const u64 m = xs(seed) & xs(seed) & xs(seed);
        if (popc((mask * m) >> 56) < 6) continue;
        memset(win, 0, sizeof(u64) << bits);
        int i;
        for (i = 0; i < n; i++) {
            const u64 idx = (occ[i] * m) >> (64 - bits);
            if (!win[idx]) win[idx] = att[i];
            else if (win[idx] != att[i]) break;
        }
        if (i == n) {
            e.mask = mask; e.factor = m; e.pos = (u32)sq << bits; e.off = 0; e.pad = 0;
            return true;
        }
    }
    return false;
}

static bool fill_fancy(int sq, bool rook, SliderTables& t, int& nuniq_r, int& nuniq_b) {
    MagicEntry& e = rook ? t.fancy_r[sq] : t.fancy_b[sq];
    const int bits = rook ? 12 : 9;
    const int n = 1 << popc(e.mask);
    u64* uniq = rook ? t.byte_r : t.byte_b;
    int& nu = rook ? nuniq_r : nuniq_b;
    const int cap = rook ? 4900 : 1428;
    e.off = (u32)nu;
    int local = 0;
    for (int i = 0; i < n; i++) {
        const u64 o = subset(e.mask, (u32)i);
        const u64 a = rook ? rook_fill(bit(sq), ~o) : bishop_fill(bit(sq), ~o);
        const u32 slot = e.pos + (u32)((o * e.factor) >> (64 - bits));
        if (slot >= 97264) return false;
        if (!t.fancy_tab[slot]) {
            t.fancy_tab[slot] = a;
            int j = 0;
            while (j < local && uniq[nu + j] != a) j++;
            if (j == local) {
                if (nu + local >= cap) return false;
                uniq[nu + local++] = a;
            }
            if (j > 255) return false;
            t.byte_tab[slot] = (u8)j;
        } else if (t.fancy_tab[slot] != a) {
            return false;
        }
    }
    nu += local;
    return true;
}

int build_slider_tables(SliderTables& t) {
    memset(&t, 0, sizeof(t));
    u64 seed = 0x2545F4914F6CDD1DULL;
    for (int sq = 0; sq < 64; sq++) {
        if (!find_plain(sq, false, seed, t.plain_b[sq], t.plain_btab + sq * 512)) return -1;
        if (!find_plain(sq, true, seed, t.plain_r[sq], t.plain_rtab + sq * 4096)) return -1;
    }
    int k = 0;
    u64 factors[128];
    u32 offsets[128];
#define FSM_ENTRY(f, o) factors[k] = (f); offsets[k] = (o); k++;
#include "fixed_shift_magics.inc"
#undef FSM_ENTRY
    if (k != 128) return -2;
    for (int sq = 0; sq < 64; sq++) {
        t.fancy_b[sq] = MagicEntry{bishop_mask(sq), factors[sq], offsets[sq], 0, 0};
        t.fancy_r[sq] = MagicEntry{rook_mask(sq), factors[64 + sq], offsets[64 + sq], 0, 0};
    }
    int nr = 0, nb = 0;
    for (int sq = 0; sq < 64; sq++) {
        if (!fill_fancy(sq, true, t, nr, nb)) return -3;
        if (!fill_fancy(sq, false, t, nr, nb)) return -4;
    }
    if (nr != 4900 || nb != 1428) return -5;
    return 0;
}

}  // namespace pf
