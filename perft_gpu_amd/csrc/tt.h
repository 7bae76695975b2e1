// tt.h -- Zobrist keys and the lockless transposition table entry format.
//
// Keys (build specification, restated by the oracle): a xorshift64(13,7,17)
// stream seeded with `seed` draws pieces[colour][type][square] (types P N B R
// Q K, colour 0 = white), castle[colour][king side, queen side], ep[file] and
// side -- 781 values; a node's key XORs the entries of its pieces (absolute
// squares), its castling rights, its en-passant file and `side` when white is
// to move (MoveGeneratorBitboard.h:3844-3880 defines the same structure over
// the reference's own random table, randoms.cpp).  Two independent key sets
// are used: key 1 (seed) picks the bucket, key 2 (seed ^ kKey2Salt) is
// stored and verified, so a false hit needs a 64-bit match on key 2 plus a
// log2(buckets)-bit match on key 1 (~95 bits at 2^31 buckets).
//
// Entry (16 bytes, replaces HashEntryPerft128b / HashKey128b of chess.h:410-504):
//   w1 = depth << 56 | count        (count < 2^56)
//   w0 = key2 ^ w1                  (lockless XOR check: a torn 16-byte write
//                                    fails verification instead of returning a
//                                    wrong count, as perft_bb.h:936-952)
// Buckets of two entries (32 bytes): slot 0 keeps the deeper subtree
// (replaced only by an equal or deeper one), slot 1 always takes the newest.
#pragma once
#include "movegen.h"

namespace pf {

constexpr u64 kKey2Salt = 0xD1B54A32D192ED03ULL;
constexpr u64 kCountMask = (1ULL << 56) - 1;

// A table's "mask word": bits 0-55 are buckets - 1; bits 56-59 hold k > 0
// when the table is laid out as 2^k equal regions, one per remaining depth
// modulo 2^k (the reference's one table per depth, launcher.h:62-126),
// instead of one table shared by every depth (k = 0, the default: the depth
// is part of each entry, so entries of all depths compete for all buckets).
PF_HD u64 tt_bucket(u64 k1, int depth, u64 mw) {
    const u64 m = mw & kCountMask;
    const int k = (int)(mw >> 56);
    if (!k) return k1 & m;
    const u64 sub = m >> k;  // buckets per region - 1
    return (u64)(depth & ((1 << k) - 1)) * (sub + 1) + (k1 & sub);
}

// The reference CPU generator's en-passant flagging (generateBoards,
// MoveGeneratorBitboard.h:1575-1582): the file is set after every double
// push, whether or not an enemy pawn can capture.  make_move flags it only
// when one can (EXACT_EN_PASSENT_FLAGGING, :3528-3547); counts are the same
// either way, keys (and so table and duplicate matches) are not.
PF_HD Board ep_on_every_double_push(Board c, u32 m) {
    if (((m >> 12) & 15) == kFlDouble) c.pawns = (c.pawns & ~0xF0ULL) | ((u64)((m & 7) + 1) << 4);
    return c;
}

struct ZKeys {
    u64 piece[2][6][64];
    u64 castle16[16];  // XOR of castle[] for each absolute rights nibble
    u64 ep[8];
    u64 side;
    u64 pad[7];
};
struct ZTables {
    ZKeys k[2];  // k[0] bucket key, k[1] verification key
};

inline void zobrist_fill(ZKeys& z, u64 seed) {
    u64 s = seed ? seed : 0x9E3779B97F4A7C15ULL;
    auto nx = [&s]() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; };
    for (int c = 0; c < 2; c++)
        for (int t = 0; t < 6; t++)
            for (int q = 0; q < 64; q++) z.piece[c][t][q] = nx();
    u64 castle[4];
    for (int i = 0; i < 4; i++) castle[i] = nx();  // white K, white Q, black K, black Q
    for (int f = 0; f < 8; f++) z.ep[f] = nx();
    z.side = nx();
    for (int r = 0; r < 16; r++) {
        u64 k = 0;
        for (int i = 0; i < 4; i++)
            if (r >> i & 1) k ^= castle[i];
        z.castle16[r] = k;
    }
    for (int i = 0; i < 7; i++) z.pad[i] = 0;
}

inline void zobrist_tables(ZTables& t, u64 seed) {
    zobrist_fill(t.k[0], seed);
    zobrist_fill(t.k[1], (seed ? seed : 0x9E3779B97F4A7C15ULL) ^ kKey2Salt);
}

// piece type index (P N B R Q K = 0..5) of an occupied square
PF_HD int type_at(const Board& b, u64 sq) {
    if (b.pawns & kRanks2to7 & sq) return 0;
    if (b.knights & sq) return 1;
    if (b.kings & sq) return 5;
    const bool bq = (b.bishops & sq) != 0, rq = (b.rooks & sq) != 0;
    return (bq && rq) ? 4 : bq ? 2 : 3;
}

// absolute rights nibble (bit0 white O-O, bit1 white O-O-O, bit2/3 black) of a
// normalised rights field (bits 0-1 side to move, 2-3 opponent)
PF_HD int abs_rights(int r, int black) { return black ? ((r >> 2) | ((r & 3) << 2)) : r; }

// Full key of a Board whose side to move has colour `black`.
PF_HD u64 zobrist_full(const ZKeys& z, const Board& b, int black) {
    const int fl = black ? 56 : 0;
    u64 k = black ? 0 : z.side;
    k ^= z.castle16[abs_rights(st_rights(b), black)];
    if (st_ep(b)) k ^= z.ep[st_ep(b) - 1];
    u64 occ = occupancy(b);
    while (occ) {
        const int s = ctz(occ);
        const u64 sb = bit(s);
        occ &= occ - 1;
        const int colour = (b.us & sb) ? black : !black;
        k ^= z.piece[colour][type_at(b, sb)][s ^ fl];
    }
    return k;
}

// Key of the child reached by legal move m from b (mover colour `black`),
// given the parent key; `child` is make_move(b, m) in its own frame.
PF_HD u64 zobrist_child(const ZKeys& z, u64 k, const Board& b, u32 m, int black, const Board& child) {
    const int from = m & 63, to = (m >> 6) & 63, fl = (m >> 12) & 15, x = black ? 56 : 0;
    const int c = black, o = !black;
    const u64 src = bit(from), dst = bit(to);
    const int p = type_at(b, src);
    const int np = (fl & 8) ? 1 + (fl & 3) : p;  // promotion N B R Q -> 1..4
    k ^= z.piece[c][p][from ^ x] ^ z.piece[c][np][to ^ x];
    if (occupancy(b) & dst) k ^= z.piece[o][type_at(b, dst)][to ^ x];
    if (fl == kFlEP) k ^= z.piece[o][0][(to - 8) ^ x];
    if (fl == kFlOO) k ^= z.piece[c][3][7 ^ x] ^ z.piece[c][3][5 ^ x];
    if (fl == kFlOOO) k ^= z.piece[c][3][0 ^ x] ^ z.piece[c][3][3 ^ x];
    k ^= z.castle16[abs_rights(st_rights(b), black)] ^ z.castle16[abs_rights(st_rights(child), !black)];
    if (st_ep(b)) k ^= z.ep[st_ep(b) - 1];
    if (st_ep(child)) k ^= z.ep[st_ep(child) - 1];
    return k ^ z.side;
}

}  // namespace pf
