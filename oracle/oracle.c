/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Plain-C restatement of the reference's CPU perft path.  Every function names
 * the reference lines it follows.  Side to move is a run-time argument here
 * (the reference templates on it); thin wrappers with a constant side let the
 * compiler specialise exactly as the template does.
 *
 * Built twice by oracle/Makefile: liboracle.so (plain) and liboracle_ops.so
 * (-DORC_OPCOUNT: every 64-bit operation on the Kogge-Stone path is tallied
 * with the frozen int32-lane-op weights: and/or/xor/not/shift/add/sub/select
 * = 2, 64-bit multiply = 4, popcount64 = 2, ctz64 = 3, table load = 0).
 */
#include "oracle.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef uint32_t u32;
typedef uint16_t u16;
typedef uint8_t u8;

#ifdef ORC_OPCOUNT
static __thread u64 g_ops __attribute__((tls_model("initial-exec")));
#define OPS(n) (g_ops += (u64)(n))
#else
#define OPS(n) ((void)0)
#endif

#define INL static inline __attribute__((always_inline))

/* ---- constants (MoveGeneratorBitboard.h:16-58) ---------------------------- */
#define RANKS2TO7 0x00FFFFFFFFFFFF00ULL
#define RANK1 0x00000000000000FFULL
#define RANK3 0x0000000000FF0000ULL
#define RANK8 0xFF00000000000000ULL
#define FILEA 0x0101010101010101ULL
#define FILEH 0x8080808080808080ULL
#define DIAG_A1H8 0x8040201008040201ULL
#define DIAG_A8H1 0x0102040810204080ULL
#define CENTRAL 0x007E7E7E7E7E7E00ULL
#define F1G1 0x60ULL
#define C1D1 0x0CULL
#define B1D1 0x0EULL
#define F8G8 0x6000000000000000ULL
#define C8D8 0x0C00000000000000ULL
#define B8D8 0x0E00000000000000ULL
#define WK_ROOK 0x0000000000000080ULL
#define WQ_ROOK 0x0000000000000001ULL
#define BK_ROOK 0x8000000000000000ULL
#define BQ_ROOK 0x0100000000000000ULL
#define BIT(i) (1ULL << (i))
#define WHITE 0
#define BLACK 1
#define CF_KING 1
#define CF_QUEEN 2

/* CMove flags (chess.h:259-293) */
#define FL_QUIET 0
#define FL_DOUBLE 1
#define FL_KCASTLE 2
#define FL_QCASTLE 3
#define FL_CAPTURE 4
#define FL_EP 5
#define FL_NPROMO 8
#define FL_BPROMO 9
#define FL_RPROMO 10
#define FL_QPROMO 11

enum { PAWN = 1, KNIGHT, BISHOP, ROOK, QUEEN, KING };

/* ---- packed state accessors (chess.h:211-226, explicit masks) ------------- */
INL int st_wc(const orc_pos *p) { return (int)(p->pawns & 3); }
INL int st_bc(const orc_pos *p) { return (int)((p->pawns >> 2) & 3); }
INL int st_ep(const orc_pos *p) { return (int)((p->pawns >> 4) & 15); }
INL int st_chance(const orc_pos *p) { return (int)(p->pawns >> 63); }
INL void set_wc(orc_pos *p, int v) { p->pawns = (p->pawns & ~3ULL) | (u64)(v & 3); }
INL void set_bc(orc_pos *p, int v) { p->pawns = (p->pawns & ~0xCULL) | ((u64)(v & 3) << 2); }
INL void set_ep(orc_pos *p, int v) { p->pawns = (p->pawns & ~0xF0ULL) | ((u64)(v & 15) << 4); }
INL void set_chance(orc_pos *p, int v) { p->pawns = (p->pawns & ~(1ULL << 63)) | ((u64)(v & 1) << 63); }
INL void set_hmc(orc_pos *p, int v) { p->pawns = (p->pawns & ~(0x7FULL << 56)) | ((u64)(v & 0x7F) << 56); }

/* ---- bit primitives (MoveGeneratorBitboard.h:63-139, 454-492, 886-899) ---- */
INL int popc(u64 x) { OPS(2); return __builtin_popcountll(x); }
INL int bscan(u64 x) { OPS(3); return __builtin_ctzll(x); }
INL u64 get_one(u64 x) { OPS(4); return x & (0 - x); }
INL int is_singular(u64 x) { OPS(6); return (x ^ (x & (0 - x))) == 0; }

INL u64 north_one(u64 x) { OPS(2); return x << 8; }
INL u64 south_one(u64 x) { OPS(2); return x >> 8; }
INL u64 east_one(u64 x) { OPS(4); return (x << 1) & ~FILEA; }
INL u64 west_one(u64 x) { OPS(4); return (x >> 1) & ~FILEH; }
INL u64 ne_one(u64 x) { OPS(4); return (x << 9) & ~FILEA; }
INL u64 nw_one(u64 x) { OPS(4); return (x << 7) & ~FILEH; }
INL u64 se_one(u64 x) { OPS(4); return (x >> 7) & ~FILEA; }
INL u64 sw_one(u64 x) { OPS(4); return (x >> 9) & ~FILEH; }

/* Kogge-Stone occluded attacks (MoveGeneratorBitboard.h:606-728) */
INL u64 north_att(u64 g, u64 p) {
    g |= (g << 8) & p; p &= (p << 8); g |= (g << 16) & p; p &= (p << 16); g |= (g << 32) & p;
    OPS(28); return g << 8;
}
INL u64 south_att(u64 g, u64 p) {
    g |= (g >> 8) & p; p &= (p >> 8); g |= (g >> 16) & p; p &= (p >> 16); g |= (g >> 32) & p;
    OPS(28); return g >> 8;
}
INL u64 east_att(u64 g, u64 p) {
    p &= ~FILEA;
    g |= (g << 1) & p; p &= (p << 1); g |= (g << 2) & p; p &= (p << 2); g |= (g << 3) & p;
    OPS(32); return (g << 1) & ~FILEA;
}
INL u64 west_att(u64 g, u64 p) {
    p &= ~FILEH;
    g |= (g >> 1) & p; p &= (p >> 1); g |= (g >> 2) & p; p &= (p >> 2); g |= (g >> 3) & p;
    OPS(32); return (g >> 1) & ~FILEH;
}
INL u64 ne_att(u64 g, u64 p) {
    p &= ~FILEA;
    g |= (g << 9) & p; p &= (p << 9); g |= (g << 18) & p; p &= (p << 18); g |= (g << 36) & p;
    OPS(32); return (g << 9) & ~FILEA;
}
INL u64 nw_att(u64 g, u64 p) {
    p &= ~FILEH;
    g |= (g << 7) & p; p &= (p << 7); g |= (g << 14) & p; p &= (p << 14); g |= (g << 28) & p;
    OPS(32); return (g << 7) & ~FILEH;
}
INL u64 se_att(u64 g, u64 p) {
    p &= ~FILEA;
    g |= (g >> 7) & p; p &= (p >> 7); g |= (g >> 14) & p; p &= (p >> 14); g |= (g >> 28) & p;
    OPS(32); return (g >> 7) & ~FILEA;
}
INL u64 sw_att(u64 g, u64 p) {
    p &= ~FILEH;
    g |= (g >> 9) & p; p &= (p >> 9); g |= (g >> 18) & p; p &= (p >> 18); g |= (g >> 36) & p;
    OPS(32); return (g >> 9) & ~FILEH;
}
/* MoveGeneratorBitboard.h:734-750 */
INL u64 bishop_ks(u64 b, u64 pro) { OPS(6); return ne_att(b, pro) | nw_att(b, pro) | se_att(b, pro) | sw_att(b, pro); }
INL u64 rook_ks(u64 r, u64 pro) { OPS(6); return north_att(r, pro) | south_att(r, pro) | east_att(r, pro) | west_att(r, pro); }

/* MoveGeneratorBitboard.h:862-880 */
INL u64 king_att(u64 k) {
    u64 a = east_one(k) | west_one(k);
    k |= a;
    a |= north_one(k) | south_one(k);
    OPS(8);
    return a;
}
INL u64 knight_att(u64 n) {
    u64 l1 = (n >> 1) & 0x7f7f7f7f7f7f7f7fULL;
    u64 l2 = (n >> 2) & 0x3f3f3f3f3f3f3f3fULL;
    u64 r1 = (n << 1) & 0xfefefefefefefefeULL;
    u64 r2 = (n << 2) & 0xfcfcfcfcfcfcfcfcULL;
    u64 h1 = l1 | r1, h2 = l2 | r2;
    OPS(2 * 8 + 2 * 2 + 2 * 7);
    return (h1 << 16) | (h1 >> 16) | (h2 << 8) | (h2 >> 8);
}

/* ---- tables (MoveGeneratorBitboard.h:982-1080; Magics.cpp) ---------------- */
static u64 T_between[64][64], T_line[64][64];
static u64 T_rook[64], T_bishop[64], T_king[64], T_knight[64];
static u64 T_rook_mask[64], T_bishop_mask[64];
static u64 T_plain_rmagic[64], T_plain_bmagic[64];
static u64 T_plain_rtab[64][4096], T_plain_btab[64][512];
static u64 T_fancy_factor[128]; /* 0..63 bishops, 64..127 rooks */
static int T_fancy_pos[128], T_fancy_off[128];
#define FSM_TABLE_ENTRIES 97264 /* size of the shared fixed-shift table */
static u64 T_fancy_tab[FSM_TABLE_ENTRIES];
static u8 T_byte_tab[FSM_TABLE_ENTRIES];
static u64 T_byte_rook[4900], T_byte_bishop[1428];
static int g_n_rook_unique, g_n_bishop_unique, g_fancy_used;
/* slider backend of the calling thread; the reference default is fancy
 * magics (switches.h:113-125).  Worker threads of orc_perft_batch inherit the
 * caller's choice explicitly. */
static __thread int g_backend __attribute__((tls_model("initial-exec"))) = ORC_FANCY;
static int g_inited;

/* MoveGeneratorBitboard.h:905-925: squaresInBetween (pure calculation) */
static u64 squares_between(int s1, int s2) {
    const u64 m1 = ~0ULL, a2a7 = 0x0001010101010100ULL, b2g7 = 0x0040201008040200ULL,
              h1b7 = 0x0002040810204080ULL;
    u64 btwn = (m1 << s1) ^ (m1 << s2);
    u64 file = (u64)((s2 & 7) - (s1 & 7));
    u64 rank = (u64)(((s2 | 7) - s1) >> 3);
    u64 line = ((file & 7) - 1) & a2a7;
    line += 2 * (((rank & 7) - 1) >> 58);
    line += (((rank - file) & 15) - 1) & b2g7;
    line += (((rank + file) & 15) - 1) & h1b7;
    line *= btwn & (0 - btwn);
    return line & btwn;
}

/* MoveGeneratorBitboard.h:928-961: squaresInLine */
static u64 squares_line(int s1, int s2) {
    int fd = (s2 & 7) - (s1 & 7), rd = ((s2 | 7) - s1) >> 3;
    int file = s1 & 7, rank = s1 >> 3;
    if (fd == 0) return FILEA << file;
    if (rd == 0) return RANK1 << (rank * 8);
    if (fd - rd == 0) {
        if (rank - file >= 0) return DIAG_A1H8 << ((rank - file) * 8);
        return DIAG_A1H8 >> ((file - rank) * 8);
    }
    if (fd + rd == 0) {
        int sh = (rank + file - 7) * 8;
        return sh >= 0 ? DIAG_A8H1 << sh : DIAG_A8H1 >> (-sh);
    }
    return 0;
}

/* Magics.cpp:27-39: getOccCombo -- i-th subset of mask (bit k of i -> k-th LSB) */
static u64 occ_combo(u64 mask, u64 i) {
    u64 op = 0;
    while (i) {
        u64 b = mask & (0 - mask);
        mask &= ~b;
        if (i & 1) op |= b;
        i >>= 1;
    }
    return op;
}

static u64 rng_state_magic;
static u64 xs64(u64 *s) {
    u64 x = *s;
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    *s = x;
    return x;
}
uint64_t orc_xorshift64(uint64_t *state) { return xs64(state); }

/* Magics.cpp:42-103 findMagicCommon, restated for a per-square private table
 * (plain magics; the reference re-searches with srand(time) on every run,
 * here the search is seeded so the tables are reproducible). */
static u64 find_plain_magic(u64 mask, int bits, u64 *table, int is_rook, int sq) {
    int nb = __builtin_popcountll(mask), nc = 1 << nb;
    static u64 occ[4096], att[4096];
    for (int i = 0; i < nc; i++) {
        occ[i] = occ_combo(mask, (u64)i);
        u64 pro = ~occ[i];
        att[i] = is_rook ? rook_ks(BIT(sq), pro) : bishop_ks(BIT(sq), pro);
    }
    for (long tries = 0; tries < 100000000L; tries++) {
        u64 m = xs64(&rng_state_magic) & xs64(&rng_state_magic) & xs64(&rng_state_magic);
        memset(table, 0, sizeof(u64) << bits);
        int i;
        for (i = 0; i < nc; i++) {
            u64 idx = (m * occ[i]) >> (64 - bits);
            if (table[idx] == 0) table[idx] = att[i];
            else if (table[idx] != att[i]) break;
        }
        if (i == nc) return m;
    }
    return 0; /* not found: orc_init reports failure */
}

/* MoveGeneratorBitboard.h:1055-1076 + Magics.cpp:42-103 with a precalculated
 * magic: fill the shared fixed-shift table and the byte-lookup tables. */
static int fill_fancy(int sq, int is_rook) {
    int slot = is_rook ? 64 + sq : sq;
    u64 mask = is_rook ? T_rook_mask[sq] : T_bishop_mask[sq];
    int bits = is_rook ? 12 : 9;
    int nb = __builtin_popcountll(mask), nc = 1 << nb;
    u64 *uniq = is_rook ? T_byte_rook : T_byte_bishop;
    int *nuniq = is_rook ? &g_n_rook_unique : &g_n_bishop_unique;
    int base = *nuniq, local = 0;
    T_fancy_off[slot] = base;
    for (int i = 0; i < nc; i++) {
        u64 o = occ_combo(mask, (u64)i);
        u64 a = is_rook ? rook_ks(BIT(sq), ~o) : bishop_ks(BIT(sq), ~o);
        u64 idx = (T_fancy_factor[slot] * o) >> (64 - bits);
        int pos = T_fancy_pos[slot] + (int)idx;
        if (pos >= FSM_TABLE_ENTRIES) return -1;
        if (T_fancy_tab[pos] == 0) {
            T_fancy_tab[pos] = a;
            int j;
            for (j = 0; j < local; j++)
                if (uniq[base + j] == a) break;
            if (j == local) {
                if (base + local >= (is_rook ? 4900 : 1428)) return -1;
                uniq[base + local] = a;
                local++;
            }
            if (j > 255) return -1;
            T_byte_tab[pos] = (u8)j;
            if (pos + 1 > g_fancy_used) g_fancy_used = pos + 1;
        } else if (T_fancy_tab[pos] != a) {
            return -1; /* destructive collision: data set is wrong */
        }
    }
    *nuniq = base + local;
    return 0;
}

int orc_init(void) {
    if (g_inited) return 0;
    for (int i = 0; i < 64; i++) {
        u64 x = BIT(i);
        T_rook[i] = north_att(x, ~0ULL) | south_att(x, ~0ULL) | east_att(x, ~0ULL) | west_att(x, ~0ULL);
        T_bishop[i] = ne_att(x, ~0ULL) | nw_att(x, ~0ULL) | se_att(x, ~0ULL) | sw_att(x, ~0ULL);
        T_knight[i] = knight_att(x);
        T_king[i] = king_att(x);
    }
    for (int i = 0; i < 64; i++)
        for (int j = 0; j < 64; j++) {
            if (i <= j) { T_between[i][j] = squares_between(i, j); T_between[j][i] = T_between[i][j]; }
            T_line[i][j] = squares_line(i, j);
        }
    /* MoveGeneratorBitboard.h:1026-1047 masks */
    for (int sq = 0; sq < 64; sq++) {
        u64 t = BIT(sq), m = T_rook[sq] & ~t;
        if (!(t & RANK1)) m &= ~RANK1;
        if (!(t & RANK8)) m &= ~RANK8;
        if (!(t & FILEA)) m &= ~FILEA;
        if (!(t & FILEH)) m &= ~FILEH;
        T_rook_mask[sq] = m;
        T_bishop_mask[sq] = T_bishop[sq] & ~t & CENTRAL;
    }
    rng_state_magic = 0x5eedULL;
    for (int sq = 0; sq < 64; sq++) {
        T_plain_rmagic[sq] = find_plain_magic(T_rook_mask[sq], 12, T_plain_rtab[sq], 1, sq);
        T_plain_bmagic[sq] = find_plain_magic(T_bishop_mask[sq], 9, T_plain_btab[sq], 0, sq);
        if (!T_plain_rmagic[sq] || !T_plain_bmagic[sq]) return -4;
    }
    {
        int k = 0;
#define FSM_ENTRY(f, o) T_fancy_factor[k] = (f); T_fancy_pos[k] = (o); k++;
        {
#include "fixed_shift_magics.h"
        }
#undef FSM_ENTRY
        if (k != 128) return -1;
    }
    memset(T_fancy_tab, 0, sizeof(T_fancy_tab));
    g_n_rook_unique = g_n_bishop_unique = 0;
    for (int sq = 0; sq < 64; sq++) {
        if (fill_fancy(sq, 1)) return -2;
        if (fill_fancy(sq, 0)) return -3;
    }
    g_inited = 1;
    return 0;
}

int orc_set_backend(int b) { int o = g_backend; if (b >= 0 && b <= 3) g_backend = b; return o; }

void orc_table_stats(int *fe, int *ru, int *bu) { *fe = g_fancy_used; *ru = g_n_rook_unique; *bu = g_n_bishop_unique; }

/* MoveGeneratorBitboard.h:753-829: single-piece magic lookups */
INL u64 magic_bishop(int sq, u64 occ_all) {
    u64 occ = occ_all & T_bishop_mask[sq];
    switch (g_backend) {
    case ORC_PLAIN: return T_plain_btab[sq][(T_plain_bmagic[sq] * occ) >> 55];
    case ORC_BYTE: {
        u64 idx = (T_fancy_factor[sq] * occ) >> 55;
        return T_byte_bishop[T_byte_tab[T_fancy_pos[sq] + idx] + T_fancy_off[sq]];
    }
    default: return T_fancy_tab[T_fancy_pos[sq] + ((T_fancy_factor[sq] * occ) >> 55)];
    }
}
INL u64 magic_rook(int sq, u64 occ_all) {
    u64 occ = occ_all & T_rook_mask[sq];
    switch (g_backend) {
    case ORC_PLAIN: return T_plain_rtab[sq][(T_plain_rmagic[sq] * occ) >> 52];
    case ORC_BYTE: {
        u64 idx = (T_fancy_factor[64 + sq] * occ) >> 52;
        return T_byte_rook[T_byte_tab[T_fancy_pos[64 + sq] + idx] + T_fancy_off[64 + sq]];
    }
    default: return T_fancy_tab[T_fancy_pos[64 + sq] + ((T_fancy_factor[64 + sq] * occ) >> 52)];
    }
}

/* bishopAttacks / rookAttacks (MoveGeneratorBitboard.h:753-829, 860-861):
 * `piece` holds exactly one bit for the magic backends; KS accepts a set. */
INL u64 bishop_attacks(u64 piece, u64 pro) {
    if (g_backend == ORC_KS) return bishop_ks(piece, pro);
    return magic_bishop(__builtin_ctzll(piece), ~pro);
}
INL u64 rook_attacks(u64 piece, u64 pro) {
    if (g_backend == ORC_KS) return rook_ks(piece, pro);
    return magic_rook(__builtin_ctzll(piece), ~pro);
}
/* multiBishopAttacks / multiRookAttacks (MoveGeneratorBitboard.h:831-856) */
INL u64 multi_bishop(u64 set, u64 pro) {
    if (g_backend == ORC_KS) return bishop_ks(set, pro);
    u64 a = 0;
    while (set) { u64 b = set & (0 - set); a |= magic_bishop(__builtin_ctzll(b), ~pro); set ^= b; }
    return a;
}
INL u64 multi_rook(u64 set, u64 pro) {
    if (g_backend == ORC_KS) return rook_ks(set, pro);
    u64 a = 0;
    while (set) { u64 b = set & (0 - set); a |= magic_rook(__builtin_ctzll(b), ~pro); set ^= b; }
    return a;
}

uint64_t orc_rook_attacks(int backend, int sq, uint64_t occ) {
    int o = g_backend; g_backend = backend;
    u64 r = rook_attacks(BIT(sq), ~occ);
    g_backend = o; return r;
}
uint64_t orc_bishop_attacks(int backend, int sq, uint64_t occ) {
    int o = g_backend; g_backend = backend;
    u64 r = bishop_attacks(BIT(sq), ~occ);
    g_backend = o; return r;
}
uint64_t orc_rook_mask(int sq) { return T_rook_mask[sq]; }
uint64_t orc_bishop_mask(int sq) { return T_bishop_mask[sq]; }

/* ---- pins and threats (MoveGeneratorBitboard.h:1213-1298) ----------------- */
INL u64 find_pinned(u64 enemyB, u64 enemyR, u64 all, int ksq) {
    u64 attackers = (T_bishop[ksq] & enemyB) | (T_rook[ksq] & enemyR);
    u64 pinned = 0;
    OPS(6);
    while (attackers) {
        u64 a = get_one(attackers);
        int asq = bscan(a);
        u64 between = T_between[asq][ksq] & all;
        OPS(2);
        if (is_singular(between)) { pinned |= between; OPS(2); }
        attackers ^= a; OPS(2);
    }
    return pinned;
}

INL u64 find_attacked(u64 empty, u64 eB, u64 eR, u64 eP, u64 eN, u64 eK, u64 myK, int enemy) {
    u64 att = 0;
    if (enemy == WHITE) att = ne_one(eP) | nw_one(eP);
    else att = se_one(eP) | sw_one(eP);
    att |= knight_att(eN);
    att |= multi_bishop(eB, empty | myK);
    att |= multi_rook(eR, empty | myK);
    att |= king_att(eK);
    OPS(14);
    return att;
}

/* ---- countMoves (MoveGeneratorBitboard.h:2828-3295) ----------------------- */
INL u32 count_out_of_check(const orc_pos *pos, u64 allPawns, u64 all, u64 mine, u64 enemy,
                           u64 pinned, u64 threatened, int ksq, int chance) {
    u32 n = 0;
    u64 king = pos->kings & mine;
    u64 attackers = 0;
    u64 eP = allPawns & enemy;
    attackers |= ((chance == WHITE) ? (ne_one(king) | nw_one(king)) : (se_one(king) | sw_one(king))) & eP;
    attackers |= knight_att(king) & (pos->knights & enemy);
    attackers |= bishop_attacks(king, ~all) & (pos->bishop_queens & enemy);
    attackers |= rook_attacks(king, ~all) & (pos->rook_queens & enemy);
    OPS(2 * 14);

    u64 km = T_king[ksq] & ~(threatened | mine);
    n += popc(km); OPS(6);

    if (is_singular(attackers)) {
        u64 safe = attackers | T_between[ksq][bscan(attackers)];
        mine &= ~pinned;
        u64 myPawns = allPawns & mine;
        u64 dblRank = RANK3 << (chance * 24);
        u64 epT = 0;
        int ep = st_ep(pos);
        OPS(12);
        if (ep) { epT = (chance == BLACK) ? BIT(ep - 1) << 16 : BIT(ep - 1) << 40; OPS(4); }
        u64 epCap = (chance == WHITE) ? south_one(epT) : north_one(epT);
        if (epCap != attackers) epT = 0;
        OPS(2);
        while (myPawns) {
            u64 pawn = get_one(myPawns);
            u64 dst = ((chance == WHITE) ? north_one(pawn) : south_one(pawn)) & ~all;
            OPS(4);
            if (dst) {
                OPS(2);
                if (dst & safe) {
                    OPS(2);
                    if (dst & (RANK1 | RANK8)) n += 4; else n++;
                } else {
                    dst = ((chance == WHITE) ? north_one(dst & dblRank) : south_one(dst & dblRank)) & safe & ~all;
                    OPS(8);
                    if (dst) n++;
                }
            }
            u64 wc = (chance == WHITE) ? nw_one(pawn) : sw_one(pawn);
            u64 ec = (chance == WHITE) ? ne_one(pawn) : se_one(pawn);
            dst = (wc | ec) & enemy & safe;
            OPS(6);
            if (dst) { OPS(2); if (dst & (RANK1 | RANK8)) n += 4; else n++; }
            dst = (wc | ec) & epT;
            OPS(4);
            if (dst) n++;
            myPawns ^= pawn; OPS(2);
        }
        u64 myN = pos->knights & mine; OPS(2);
        while (myN) {
            u64 kn = get_one(myN);
            n += popc(T_knight[bscan(kn)] & safe); OPS(2);
            myN ^= kn; OPS(2);
        }
        u64 bs = pos->bishop_queens & mine; OPS(2);
        while (bs) {
            u64 b = get_one(bs);
            n += popc(bishop_attacks(b, ~all) & safe); OPS(4);
            bs ^= b; OPS(2);
        }
        u64 rs = pos->rook_queens & mine; OPS(2);
        while (rs) {
            u64 r = get_one(rs);
            n += popc(rook_attacks(r, ~all) & safe); OPS(4);
            rs ^= r; OPS(2);
        }
    }
    return n;
}

INL u32 count_moves_c(const orc_pos *pos, int chance) {
    u32 n = 0;
    u64 allPawns = pos->pawns & RANKS2TO7;
    u64 all = pos->kings | allPawns | pos->knights | pos->bishop_queens | pos->rook_queens;
    u64 black = all & ~pos->white;
    u64 mine = (chance == WHITE) ? pos->white : black;
    u64 enemy = (chance == WHITE) ? black : pos->white;
    u64 eB = pos->bishop_queens & enemy, eR = pos->rook_queens & enemy;
    u64 myK = pos->kings & mine;
    OPS(2 * 12);
    int ksq = bscan(myK);
    u64 pinned = find_pinned(eB, eR, all, ksq);
    u64 threatened = find_attacked(~all, eB, eR, allPawns & enemy, pos->knights & enemy,
                                   pos->kings & enemy, myK, !chance);
    OPS(10);
    if (threatened & myK)
        return count_out_of_check(pos, allPawns, all, mine, enemy, pinned, threatened, ksq, chance);

    u64 myPawns = allPawns & mine;
    u64 epT = 0;
    int ep = st_ep(pos);
    OPS(4);
    if (ep) { epT = (chance == BLACK) ? BIT(ep - 1) << 16 : BIT(ep - 1) << 40; OPS(4); }
    if (epT) {
        u64 epCap = (chance == WHITE) ? south_one(epT) : north_one(epT);
        u64 src = (east_one(epCap) | west_one(epCap)) & myPawns;
        OPS(4);
        while (src) {
            u64 pawn = get_one(src);
            OPS(2);
            if (pawn & pinned) {
                u64 line = T_line[bscan(pawn)][ksq];
                OPS(2);
                if (epT & line) n++;
            } else {
                u64 pro = ~all | epCap | pawn;
                u64 chk = (east_att(eR, pro) | west_att(eR, pro)) & myK;
                OPS(10);
                if (!chk) n++;
            }
            src ^= pawn; OPS(2);
        }
    }
    u64 dblRank = RANK3 << (chance * 24);
    u64 pp = myPawns & pinned;
    OPS(4);
    while (pp) {
        u64 pawn = get_one(pp);
        int psq = bscan(pawn);
        u64 line = T_line[psq][ksq];
        u64 dst = ((chance == WHITE) ? north_one(pawn) : south_one(pawn)) & line & ~all;
        OPS(6);
        if (dst) {
            n++;
            dst = ((chance == WHITE) ? north_one(dst & dblRank) : south_one(dst & dblRank)) & ~all;
            OPS(6);
            if (dst) n++;
        }
        dst = ((chance == WHITE) ? nw_one(pawn) : sw_one(pawn)) & line;
        dst |= ((chance == WHITE) ? ne_one(pawn) : se_one(pawn)) & line;
        OPS(8);
        if (dst & enemy) { OPS(2); if (dst & (RANK1 | RANK8)) n += 4; else n++; }
        pp ^= pawn; OPS(2);
    }
    myPawns &= ~pinned;
    u64 d = ((chance == WHITE) ? north_one(myPawns) : south_one(myPawns)) & ~all;
    n += popc(d);
    n += 3 * popc(d & (RANK1 | RANK8));
    d = ((chance == WHITE) ? north_one(d & dblRank) : south_one(d & dblRank)) & ~all;
    n += popc(d);
    d = ((chance == WHITE) ? nw_one(myPawns) : sw_one(myPawns)) & enemy;
    n += popc(d);
    n += 3 * popc(d & (RANK1 | RANK8));
    d = ((chance == WHITE) ? ne_one(myPawns) : se_one(myPawns)) & enemy;
    n += popc(d);
    n += 3 * popc(d & (RANK1 | RANK8));
    OPS(2 * 14);

    if (chance == WHITE) {
        int wc = st_wc(pos);
        if ((wc & CF_KING) && !(F1G1 & all) && !(F1G1 & threatened)) n++;
        if ((wc & CF_QUEEN) && !(B1D1 & all) && !(C1D1 & threatened)) n++;
    } else {
        int bc = st_bc(pos);
        if ((bc & CF_KING) && !(F8G8 & all) && !(F8G8 & threatened)) n++;
        if ((bc & CF_QUEEN) && !(B8D8 & all) && !(C8D8 & threatened)) n++;
    }
    OPS(2 * 8);
    u64 km = T_king[ksq] & ~(threatened | mine);
    n += popc(km); OPS(6);

    u64 myN = pos->knights & mine & ~pinned; OPS(4);
    while (myN) {
        u64 kn = get_one(myN);
        n += popc(T_knight[bscan(kn)] & ~mine); OPS(4);
        myN ^= kn; OPS(2);
    }
    u64 myB = pos->bishop_queens & mine;
    u64 bs = myB & pinned; OPS(4);
    while (bs) {
        u64 b = get_one(bs);
        u64 mv = bishop_attacks(b, ~all) & ~mine;
        mv &= T_line[bscan(b)][ksq];
        n += popc(mv); OPS(8);
        bs ^= b; OPS(2);
    }
    bs = myB & ~pinned; OPS(2);
    while (bs) {
        u64 b = get_one(bs);
        n += popc(bishop_attacks(b, ~all) & ~mine); OPS(6);
        bs ^= b; OPS(2);
    }
    u64 myR = pos->rook_queens & mine;
    u64 rs = myR & pinned; OPS(4);
    while (rs) {
        u64 r = get_one(rs);
        u64 mv = rook_attacks(r, ~all) & ~mine;
        mv &= T_line[bscan(r)][ksq];
        n += popc(mv); OPS(8);
        rs ^= r; OPS(2);
    }
    rs = myR & ~pinned; OPS(2);
    while (rs) {
        u64 r = get_one(rs);
        n += popc(rook_attacks(r, ~all) & ~mine); OPS(6);
        rs ^= r; OPS(2);
    }
    return n;
}

static u32 count_w(const orc_pos *p) { return count_moves_c(p, WHITE); }
static u32 count_b(const orc_pos *p) { return count_moves_c(p, BLACK); }
uint32_t orc_count_moves(const orc_pos *p) { return st_chance(p) ? count_b(p) : count_w(p); }

/* ---- generation (MoveGeneratorBitboard.h:1302-2823) ----------------------- */
/* One traversal in generateBoards order; the sink either writes child boards
 * (add*Move, MoveGeneratorBitboard.h:1302-1740) or CMoves (addCompactMove,
 * MoveGeneratorBitboard.h:1700-1740). */
typedef struct {
    int mode; /* 0 boards, 1 moves */
    orc_pos *boards;
    u16 *moves;
    u32 n;
} sink_t;

INL void emit_move(sink_t *s, int from, int to, int flags) {
    s->moves[s->n++] = (u16)(((flags & 0xF) << 12) | ((to & 0x3F) << 6) | (from & 0x3F));
}

/* updateCastleFlag (MoveGeneratorBitboard.h:1310-1340), if/else form */
INL void update_castle(orc_pos *p, u64 dst, int chance) {
    if (chance == WHITE) {
        if (dst & BK_ROOK) set_bc(p, st_bc(p) & ~CF_KING);
        else if (dst & BQ_ROOK) set_bc(p, st_bc(p) & ~CF_QUEEN);
    } else {
        if (dst & WK_ROOK) set_wc(p, st_wc(p) & ~CF_KING);
        else if (dst & WQ_ROOK) set_wc(p, st_wc(p) & ~CF_QUEEN);
    }
}

/* addSlidingMove (MoveGeneratorBitboard.h:1342-1400) */
INL void add_sliding(sink_t *s, const orc_pos *pos, u64 src, u64 dst, int chance, u64 enemy) {
    if (s->mode) { emit_move(s, __builtin_ctzll(src), __builtin_ctzll(dst), (dst & enemy) ? FL_CAPTURE : 0); return; }
    orc_pos nb;
    nb.bishop_queens = pos->bishop_queens & ~dst;
    nb.rook_queens = pos->rook_queens & ~dst;
    nb.kings = pos->kings & ~dst;
    nb.knights = pos->knights & ~dst;
    nb.pawns = pos->pawns & ~(dst & RANKS2TO7);
    u64 isB = nb.bishop_queens & src, isR = nb.rook_queens & src;
    nb.bishop_queens ^= isB; nb.rook_queens ^= isR;
    nb.bishop_queens |= isB ? dst : 0;
    nb.rook_queens |= isR ? dst : 0;
    nb.white = (chance == WHITE) ? ((pos->white ^ src) | dst) : (pos->white & ~dst);
    set_chance(&nb, !chance);
    set_ep(&nb, 0);
    update_castle(&nb, dst, chance);
    update_castle(&nb, src, !chance);
    s->boards[s->n++] = nb;
}

/* addKnightMove (MoveGeneratorBitboard.h:1403-1443) */
INL void add_knight(sink_t *s, const orc_pos *pos, u64 src, u64 dst, int chance, u64 enemy) {
    if (s->mode) { emit_move(s, __builtin_ctzll(src), __builtin_ctzll(dst), (dst & enemy) ? FL_CAPTURE : 0); return; }
    orc_pos nb;
    nb.bishop_queens = pos->bishop_queens & ~dst;
    nb.rook_queens = pos->rook_queens & ~dst;
    nb.kings = pos->kings & ~dst;
    nb.pawns = pos->pawns & ~(dst & RANKS2TO7);
    nb.knights = (pos->knights ^ src) | dst;
    nb.white = (chance == WHITE) ? ((pos->white ^ src) | dst) : (pos->white & ~dst);
    set_chance(&nb, !chance);
    set_ep(&nb, 0);
    update_castle(&nb, dst, chance);
    s->boards[s->n++] = nb;
}

/* addKingMove (MoveGeneratorBitboard.h:1446-1488) */
INL void add_king(sink_t *s, const orc_pos *pos, u64 src, u64 dst, int chance, u64 enemy) {
    if (s->mode) { emit_move(s, __builtin_ctzll(src), __builtin_ctzll(dst), (dst & enemy) ? FL_CAPTURE : 0); return; }
    orc_pos nb;
    nb.bishop_queens = pos->bishop_queens & ~dst;
    nb.rook_queens = pos->rook_queens & ~dst;
    nb.knights = pos->knights & ~dst;
    nb.pawns = pos->pawns & ~(dst & RANKS2TO7);
    nb.kings = (pos->kings ^ src) | dst;
    if (chance == WHITE) { nb.white = (pos->white ^ src) | dst; set_wc(&nb, 0); }
    else { nb.white = pos->white & ~dst; set_bc(&nb, 0); }
    set_chance(&nb, !chance);
    set_ep(&nb, 0);
    update_castle(&nb, dst, chance);
    s->boards[s->n++] = nb;
}

/* addCastleMove (MoveGeneratorBitboard.h:1491-1528) */
INL void add_castle(sink_t *s, const orc_pos *pos, int kf, int kt, int rf, int rt, int chance, int flag) {
    if (s->mode) { emit_move(s, kf, kt, flag); return; }
    orc_pos nb;
    nb.bishop_queens = pos->bishop_queens;
    nb.pawns = pos->pawns;
    nb.knights = pos->knights;
    nb.kings = (pos->kings ^ BIT(kf)) | BIT(kt);
    nb.rook_queens = (pos->rook_queens ^ BIT(rf)) | BIT(rt);
    set_chance(&nb, !chance);
    set_ep(&nb, 0);
    set_hmc(&nb, 0);
    if (chance == WHITE) { nb.white = (pos->white ^ (BIT(kf) | BIT(rf))) | (BIT(kt) | BIT(rt)); set_wc(&nb, 0); }
    else { set_bc(&nb, 0); nb.white = pos->white; }
    s->boards[s->n++] = nb;
}

/* addSinglePawnMove (MoveGeneratorBitboard.h:1533-1580) */
INL void add_single_pawn(sink_t *s, const orc_pos *pos, u64 src, u64 dst, int chance, int dbl, int psq, u64 enemy) {
    if (s->mode) {
        emit_move(s, __builtin_ctzll(src), __builtin_ctzll(dst), dbl ? FL_DOUBLE : ((dst & enemy) ? FL_CAPTURE : 0));
        return;
    }
    orc_pos nb;
    nb.bishop_queens = pos->bishop_queens & ~dst;
    nb.rook_queens = pos->rook_queens & ~dst;
    nb.knights = pos->knights & ~dst;
    nb.kings = pos->kings;
    nb.pawns = (pos->pawns ^ src) | dst;
    nb.white = (chance == WHITE) ? ((pos->white ^ src) | dst) : (pos->white & ~dst);
    set_chance(&nb, !chance);
    set_ep(&nb, dbl ? (psq & 7) + 1 : 0);
    set_hmc(&nb, 0);
    s->boards[s->n++] = nb;
}

/* addEnPassentMove (MoveGeneratorBitboard.h:1582-1625) */
INL void add_ep(sink_t *s, const orc_pos *pos, u64 src, u64 dst, int chance) {
    if (s->mode) { emit_move(s, __builtin_ctzll(src), __builtin_ctzll(dst), FL_EP); return; }
    orc_pos nb;
    u64 cap = (chance == WHITE) ? (dst >> 8) : (dst << 8);
    nb.bishop_queens = pos->bishop_queens;
    nb.rook_queens = pos->rook_queens;
    nb.knights = pos->knights;
    nb.kings = pos->kings;
    nb.pawns = (pos->pawns ^ (cap | src)) | dst;
    nb.white = (chance == WHITE) ? ((pos->white ^ src) | dst) : (pos->white ^ cap);
    set_chance(&nb, !chance);
    set_hmc(&nb, 0);
    set_ep(&nb, 0);
    s->boards[s->n++] = nb;
}

/* addPawnMoves (MoveGeneratorBitboard.h:1629-1697): promotions N, B, Q, R */
INL void add_pawn_moves(sink_t *s, const orc_pos *pos, u64 src, u64 dst, int chance, u64 enemy) {
    if (dst & (RANK1 | RANK8)) {
        if (s->mode) {
            int f = __builtin_ctzll(src), t = __builtin_ctzll(dst), c = (dst & enemy) ? FL_CAPTURE : 0;
            emit_move(s, f, t, c | FL_NPROMO);
            emit_move(s, f, t, c | FL_BPROMO);
            emit_move(s, f, t, c | FL_QPROMO);
            emit_move(s, f, t, c | FL_RPROMO);
            return;
        }
        orc_pos nb;
        nb.kings = pos->kings;
        nb.white = (chance == WHITE) ? ((pos->white ^ src) | dst) : (pos->white & ~dst);
        nb.pawns = pos->pawns ^ src;
        set_chance(&nb, !chance);
        set_ep(&nb, 0);
        set_hmc(&nb, 0);
        update_castle(&nb, dst, chance);
        nb.knights = pos->knights | dst;
        nb.bishop_queens = pos->bishop_queens & ~dst;
        nb.rook_queens = pos->rook_queens & ~dst;
        s->boards[s->n++] = nb;
        nb.knights = pos->knights & ~dst;
        nb.bishop_queens = pos->bishop_queens | dst;
        nb.rook_queens = pos->rook_queens & ~dst;
        s->boards[s->n++] = nb;
        nb.rook_queens = pos->rook_queens | dst;
        s->boards[s->n++] = nb;
        nb.bishop_queens = pos->bishop_queens & ~dst;
        s->boards[s->n++] = nb;
    } else {
        add_single_pawn(s, pos, src, dst, chance, 0, 0, enemy);
    }
}

/* generateBoardsOutOfCheck (MoveGeneratorBitboard.h:1744-1943) */
INL void gen_out_of_check(sink_t *s, const orc_pos *pos, u64 allPawns, u64 all, u64 mine, u64 enemy,
                          u64 pinned, u64 threatened, int ksq, int chance) {
    u64 king = pos->kings & mine;
    u64 attackers = 0;
    u64 eP = allPawns & enemy;
    attackers |= ((chance == WHITE) ? (ne_one(king) | nw_one(king)) : (se_one(king) | sw_one(king))) & eP;
    attackers |= knight_att(king) & (pos->knights & enemy);
    attackers |= bishop_attacks(king, ~all) & (pos->bishop_queens & enemy);
    attackers |= rook_attacks(king, ~all) & (pos->rook_queens & enemy);
    OPS(2 * 14);
    u64 km = T_king[ksq] & ~(threatened | mine);
    OPS(4);
    while (km) { u64 d = get_one(km); add_king(s, pos, king, d, chance, enemy); km ^= d; OPS(2 + 12); }
    if (!is_singular(attackers)) return;
    u64 safe = attackers | T_between[ksq][bscan(attackers)];
    mine &= ~pinned;
    u64 myPawns = allPawns & mine;
    u64 dblRank = RANK3 << (chance * 24);
    u64 epT = 0;
    int ep = st_ep(pos);
    OPS(12);
    if (ep) epT = (chance == BLACK) ? BIT(ep - 1) << 16 : BIT(ep - 1) << 40;
    u64 epCap = (chance == WHITE) ? south_one(epT) : north_one(epT);
    if (epCap != attackers) epT = 0;
    while (myPawns) {
        u64 pawn = get_one(myPawns);
        u64 dst = ((chance == WHITE) ? north_one(pawn) : south_one(pawn)) & ~all;
        OPS(4);
        if (dst) {
            if (dst & safe) add_pawn_moves(s, pos, pawn, dst, chance, enemy);
            else {
                dst = ((chance == WHITE) ? north_one(dst & dblRank) : south_one(dst & dblRank)) & safe & ~all;
                OPS(8);
                if (dst) add_single_pawn(s, pos, pawn, dst, chance, 1, __builtin_ctzll(pawn), enemy);
            }
        }
        u64 wc = (chance == WHITE) ? nw_one(pawn) : sw_one(pawn);
        u64 ec = (chance == WHITE) ? ne_one(pawn) : se_one(pawn);
        dst = (wc | ec) & enemy & safe;
        OPS(6);
        if (dst) add_pawn_moves(s, pos, pawn, dst, chance, enemy);
        dst = (wc | ec) & epT;
        OPS(4);
        if (dst) add_ep(s, pos, pawn, dst, chance);
        myPawns ^= pawn; OPS(2);
    }
    u64 myN = pos->knights & mine;
    while (myN) {
        u64 kn = get_one(myN);
        u64 mv = T_knight[bscan(kn)] & safe;
        while (mv) { u64 d = get_one(mv); add_knight(s, pos, kn, d, chance, enemy); mv ^= d; OPS(2 + 12); }
        myN ^= kn; OPS(4);
    }
    u64 bs = pos->bishop_queens & mine;
    while (bs) {
        u64 b = get_one(bs);
        u64 mv = bishop_attacks(b, ~all) & safe;
        while (mv) { u64 d = get_one(mv); add_sliding(s, pos, b, d, chance, enemy); mv ^= d; OPS(2 + 12); }
        bs ^= b; OPS(6);
    }
    u64 rs = pos->rook_queens & mine;
    while (rs) {
        u64 r = get_one(rs);
        u64 mv = rook_attacks(r, ~all) & safe;
        while (mv) { u64 d = get_one(mv); add_sliding(s, pos, r, d, chance, enemy); mv ^= d; OPS(2 + 12); }
        rs ^= r; OPS(6);
    }
}

/* generateBoards (MoveGeneratorBitboard.h:1950-2261) */
INL void gen_c(sink_t *s, const orc_pos *pos, int chance) {
    u64 allPawns = pos->pawns & RANKS2TO7;
    u64 all = pos->kings | allPawns | pos->knights | pos->bishop_queens | pos->rook_queens;
    u64 black = all & ~pos->white;
    u64 mine = (chance == WHITE) ? pos->white : black;
    u64 enemy = (chance == WHITE) ? black : pos->white;
    u64 eB = pos->bishop_queens & enemy, eR = pos->rook_queens & enemy;
    u64 myK = pos->kings & mine;
    OPS(2 * 12);
    int ksq = bscan(myK);
    u64 pinned = find_pinned(eB, eR, all, ksq);
    u64 threatened = find_attacked(~all, eB, eR, allPawns & enemy, pos->knights & enemy, pos->kings & enemy, myK, !chance);
    OPS(10);
    if (threatened & myK) {
        gen_out_of_check(s, pos, allPawns, all, mine, enemy, pinned, threatened, ksq, chance);
        return;
    }
    u64 myPawns = allPawns & mine;
    u64 epT = 0;
    int ep = st_ep(pos);
    OPS(4);
    if (ep) epT = (chance == BLACK) ? BIT(ep - 1) << 16 : BIT(ep - 1) << 40;
    if (epT) {
        u64 epCap = (chance == WHITE) ? south_one(epT) : north_one(epT);
        u64 src = (east_one(epCap) | west_one(epCap)) & myPawns;
        OPS(4);
        while (src) {
            u64 pawn = get_one(src);
            if (pawn & pinned) {
                u64 line = T_line[bscan(pawn)][ksq];
                if (epT & line) add_ep(s, pos, pawn, epT, chance);
            } else {
                u64 pro = ~all | epCap | pawn;
                u64 chk = (east_att(eR, pro) | west_att(eR, pro)) & myK;
                OPS(10);
                if (!chk) add_ep(s, pos, pawn, epT, chance);
            }
            src ^= pawn; OPS(6);
        }
    }
    u64 dblRank = RANK3 << (chance * 24);
    u64 pp = myPawns & pinned;
    OPS(4);
    while (pp) {
        u64 pawn = get_one(pp);
        int psq = bscan(pawn);
        u64 line = T_line[psq][ksq];
        u64 dst = ((chance == WHITE) ? north_one(pawn) : south_one(pawn)) & line & ~all;
        OPS(6);
        if (dst) {
            add_single_pawn(s, pos, pawn, dst, chance, 0, psq, enemy);
            dst = ((chance == WHITE) ? north_one(dst & dblRank) : south_one(dst & dblRank)) & ~all;
            OPS(6);
            if (dst) add_single_pawn(s, pos, pawn, dst, chance, 1, psq, enemy);
        }
        dst = ((chance == WHITE) ? nw_one(pawn) : sw_one(pawn)) & line;
        dst |= ((chance == WHITE) ? ne_one(pawn) : se_one(pawn)) & line;
        OPS(8);
        if (dst & enemy) add_pawn_moves(s, pos, pawn, dst, chance, enemy);
        pp ^= pawn; OPS(2);
    }
    myPawns &= ~pinned;
    OPS(2);
    while (myPawns) {
        u64 pawn = get_one(myPawns);
        u64 dst = ((chance == WHITE) ? north_one(pawn) : south_one(pawn)) & ~all;
        OPS(2);
        if (dst) {
            add_pawn_moves(s, pos, pawn, dst, chance, enemy);
            dst = ((chance == WHITE) ? north_one(dst & dblRank) : south_one(dst & dblRank)) & ~all;
            OPS(6);
            if (dst) add_single_pawn(s, pos, pawn, dst, chance, 1, __builtin_ctzll(pawn), enemy);
        }
        u64 wc = (chance == WHITE) ? nw_one(pawn) : sw_one(pawn);
        dst = wc & enemy;
        OPS(2);
        if (dst) add_pawn_moves(s, pos, pawn, dst, chance, enemy);
        u64 ec = (chance == WHITE) ? ne_one(pawn) : se_one(pawn);
        dst = ec & enemy;
        OPS(2);
        if (dst) add_pawn_moves(s, pos, pawn, dst, chance, enemy);
        myPawns ^= pawn; OPS(2);
    }
    if (chance == WHITE) {
        int wc = st_wc(pos);
        if ((wc & CF_KING) && !(F1G1 & all) && !(F1G1 & threatened)) add_castle(s, pos, 4, 6, 7, 5, chance, FL_KCASTLE);
        if ((wc & CF_QUEEN) && !(B1D1 & all) && !(C1D1 & threatened)) add_castle(s, pos, 4, 2, 0, 3, chance, FL_QCASTLE);
    } else {
        int bc = st_bc(pos);
        if ((bc & CF_KING) && !(F8G8 & all) && !(F8G8 & threatened)) add_castle(s, pos, 60, 62, 63, 61, chance, FL_KCASTLE);
        if ((bc & CF_QUEEN) && !(B8D8 & all) && !(C8D8 & threatened)) add_castle(s, pos, 60, 58, 56, 59, chance, FL_QCASTLE);
    }
    OPS(2 * 8);
    u64 km = T_king[ksq] & ~(threatened | mine);
    OPS(4);
    while (km) { u64 d = get_one(km); add_king(s, pos, myK, d, chance, enemy); km ^= d; OPS(2 + 12); }
    u64 myN = pos->knights & mine & ~pinned;
    OPS(4);
    while (myN) {
        u64 kn = get_one(myN);
        u64 mv = T_knight[bscan(kn)] & ~mine;
        while (mv) { u64 d = get_one(mv); add_knight(s, pos, kn, d, chance, enemy); mv ^= d; OPS(2 + 12); }
        myN ^= kn; OPS(4);
    }
    u64 myB = pos->bishop_queens & mine;
    u64 bs = myB & pinned;
    OPS(4);
    while (bs) {
        u64 b = get_one(bs);
        u64 mv = bishop_attacks(b, ~all) & ~mine;
        mv &= T_line[bscan(b)][ksq];
        while (mv) { u64 d = get_one(mv); add_sliding(s, pos, b, d, chance, enemy); mv ^= d; OPS(2 + 12); }
        bs ^= b; OPS(8);
    }
    bs = myB & ~pinned;
    OPS(2);
    while (bs) {
        u64 b = get_one(bs);
        u64 mv = bishop_attacks(b, ~all) & ~mine;
        while (mv) { u64 d = get_one(mv); add_sliding(s, pos, b, d, chance, enemy); mv ^= d; OPS(2 + 12); }
        bs ^= b; OPS(6);
    }
    u64 myR = pos->rook_queens & mine;
    u64 rs = myR & pinned;
    OPS(4);
    while (rs) {
        u64 r = get_one(rs);
        u64 mv = rook_attacks(r, ~all) & ~mine;
        mv &= T_line[bscan(r)][ksq];
        while (mv) { u64 d = get_one(mv); add_sliding(s, pos, r, d, chance, enemy); mv ^= d; OPS(2 + 12); }
        rs ^= r; OPS(8);
    }
    rs = myR & ~pinned;
    OPS(2);
    while (rs) {
        u64 r = get_one(rs);
        u64 mv = rook_attacks(r, ~all) & ~mine;
        while (mv) { u64 d = get_one(mv); add_sliding(s, pos, r, d, chance, enemy); mv ^= d; OPS(2 + 12); }
        rs ^= r; OPS(6);
    }
}

static u32 gen_boards_w(const orc_pos *p, orc_pos *out) { sink_t s = {0, out, 0, 0}; gen_c(&s, p, WHITE); return s.n; }
static u32 gen_boards_b(const orc_pos *p, orc_pos *out) { sink_t s = {0, out, 0, 0}; gen_c(&s, p, BLACK); return s.n; }
uint32_t orc_generate_boards(const orc_pos *p, orc_pos *out) { return st_chance(p) ? gen_boards_b(p, out) : gen_boards_w(p, out); }

static u32 gen_moves_w(const orc_pos *p, u16 *out) { sink_t s = {1, 0, out, 0}; gen_c(&s, p, WHITE); return s.n; }
static u32 gen_moves_b(const orc_pos *p, u16 *out) { sink_t s = {1, 0, out, 0}; gen_c(&s, p, BLACK); return s.n; }
uint32_t orc_generate_moves(const orc_pos *p, uint16_t *out) { return st_chance(p) ? gen_moves_b(p, out) : gen_moves_w(p, out); }

int orc_in_check(const orc_pos *pos) {
    int chance = st_chance(pos);
    u64 allPawns = pos->pawns & RANKS2TO7;
    u64 all = pos->kings | allPawns | pos->knights | pos->bishop_queens | pos->rook_queens;
    u64 black = all & ~pos->white;
    u64 mine = (chance == WHITE) ? pos->white : black;
    u64 enemy = (chance == WHITE) ? black : pos->white;
    u64 myK = pos->kings & mine;
    int o = g_backend; g_backend = ORC_KS;
    u64 th = find_attacked(~all, pos->bishop_queens & enemy, pos->rook_queens & enemy, allPawns & enemy,
                           pos->knights & enemy, pos->kings & enemy, myK, !chance);
    g_backend = o;
    return (th & myK) != 0;
}

/* ---- makeMove (MoveGeneratorBitboard.h:3310-3560, updateHash = false) ----- */
INL void make_move_c(orc_pos *pos, u16 move, int chance, int exact_ep) {
    int from = move & 63, to = (move >> 6) & 63, flags = move >> 12;
    u64 src = BIT(from), dst = BIT(to);
    u64 queens = pos->bishop_queens & pos->rook_queens;
    int piece;
    OPS(2 + 4 + 4);
    if (pos->kings & src) piece = KING;
    else if (pos->knights & src) piece = KNIGHT;
    else if ((pos->pawns & RANKS2TO7) & src) piece = PAWN;
    else if (queens & src) piece = QUEEN;
    else if (pos->bishop_queens & src) piece = BISHOP;
    else piece = ROOK;
    OPS(2 * 7);
    if (flags == 8 || flags == 12) piece = KNIGHT;
    else if (flags == 9 || flags == 13) piece = BISHOP;
    else if (flags == 10 || flags == 14) piece = ROOK;
    else if (flags == 11 || flags == 15) piece = QUEEN;
    OPS(2 * 4);
    pos->bishop_queens &= ~src; pos->rook_queens &= ~src; pos->kings &= ~src; pos->knights &= ~src;
    pos->pawns &= ~(src & RANKS2TO7);
    pos->bishop_queens &= ~dst; pos->rook_queens &= ~dst; pos->kings &= ~dst; pos->knights &= ~dst;
    pos->pawns &= ~(dst & RANKS2TO7);
    OPS(2 * 14);
    if (piece == KING) { pos->kings |= dst; if (chance == WHITE) set_wc(pos, 0); else set_bc(pos, 0); }
    if (piece == KNIGHT) pos->knights |= dst;
    if (piece == PAWN) pos->pawns |= dst;
    if (piece == BISHOP || piece == QUEEN) pos->bishop_queens |= dst;
    if (piece == ROOK || piece == QUEEN) pos->rook_queens |= dst;
    OPS(2 * 8);
    if (chance == WHITE) pos->white = (pos->white ^ src) | dst;
    else pos->white &= ~dst;
    OPS(4);
    if (flags == FL_EP) {
        u64 cap = (chance == WHITE) ? (dst >> 8) : (dst << 8);
        pos->pawns &= ~(cap & RANKS2TO7);
        if (chance == BLACK) pos->white &= ~cap;
        OPS(8);
    }
    if (chance == WHITE) {
        if (flags == FL_KCASTLE) { pos->rook_queens = (pos->rook_queens ^ BIT(7)) | BIT(5); pos->white = (pos->white ^ BIT(7)) | BIT(5); }
        else if (flags == FL_QCASTLE) { pos->rook_queens = (pos->rook_queens ^ BIT(0)) | BIT(3); pos->white = (pos->white ^ BIT(0)) | BIT(3); }
    } else {
        if (flags == FL_KCASTLE) pos->rook_queens = (pos->rook_queens ^ BIT(63)) | BIT(61);
        else if (flags == FL_QCASTLE) pos->rook_queens = (pos->rook_queens ^ BIT(56)) | BIT(59);
    }
    OPS(2 * 6);
    set_chance(pos, !chance);
    set_ep(pos, 0);
    update_castle(pos, dst, chance);
    if (piece == ROOK) update_castle(pos, src, !chance);
    OPS(2 * 10);
    if (flags == FL_DOUBLE) {
        int set = 1;
        if (exact_ep) {
            /* EXACT_EN_PASSENT_FLAGGING (MoveGeneratorBitboard.h:3528-3547) */
            u64 allPawns = pos->pawns & RANKS2TO7;
            u64 all = pos->kings | allPawns | pos->knights | pos->bishop_queens | pos->rook_queens;
            u64 black = all & ~pos->white;
            u64 enemy = (chance == WHITE) ? black : pos->white;
            u64 eP = allPawns & enemy;
            set = ((east_one(dst) | west_one(dst)) & eP) != 0;
            OPS(2 * 10);
        }
        if (set) set_ep(pos, (from & 7) + 1);
    }
}

void orc_make_move(orc_pos *p, uint16_t m, int exact_ep) {
    if (st_chance(p)) make_move_c(p, m, BLACK, exact_ep);
    else make_move_c(p, m, WHITE, exact_ep);
}

/* ---- perft_bb (perft_bb.h:79-101) ----------------------------------------- */
static u64 perft_rec(const orc_pos *pos, int depth) {
    if (depth == 1) return orc_count_moves(pos);
    orc_pos kids[256];
    u32 n = orc_generate_boards(pos, kids);
    u64 c = 0;
    for (u32 i = 0; i < n; i++) c += perft_rec(&kids[i], depth - 1);
    return c;
}

uint64_t orc_perft(const orc_pos *p, int depth) {
    if (depth <= 0) return 1;
    return perft_rec(p, depth);
}

/* Positions are handed out one at a time from a shared counter, so that
 * subtrees of very different sizes keep every thread busy to the end. */
typedef struct { const orc_pos *p; int n, depth, backend; int *next; uint64_t *out; } job_t;
static void *perft_worker(void *arg) {
    job_t *j = (job_t *)arg;
    g_backend = j->backend;
    for (;;) {
        const int i = __atomic_fetch_add(j->next, 1, __ATOMIC_RELAXED);
        if (i >= j->n) break;
        j->out[i] = orc_perft(&j->p[i], j->depth);
    }
    return 0;
}

void orc_perft_batch(const orc_pos *p, int n, int depth, int nthreads, uint64_t *out) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    job_t jobs[256];
    int next = 0;
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = (job_t){p, n, depth, g_backend, &next, out};
        pthread_create(&th[t], 0, perft_worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], 0);
}

int orc_divide(const orc_pos *p, int depth, uint16_t *moves, uint64_t *counts) {
    u16 mv[256];
    u32 n = orc_generate_moves(p, mv);
    for (u32 i = 0; i < n; i++) {
        orc_pos c = *p;
        orc_make_move(&c, mv[i], 0);
        moves[i] = mv[i];
        counts[i] = depth <= 1 ? 1 : orc_perft(&c, depth - 1);
    }
    return (int)n;
}

/* ---- FEN (util.cpp:189-280 and util.cpp:92-140) --------------------------- */
int orc_fen_to_pos(const char *fen, orc_pos *out) {
    memset(out, 0, sizeof(*out));
    int i = 0, row = 0, col = 0, placed = 0;
    for (; fen[i]; i++) {
        char c = fen[i];
        if (c == '/' || c == '\\') { row++; col = 0; }
        else if (c >= '1' && c <= '8') col += c - '0';
        else {
            const char *pcs = "PNBRQKpnbrqk";
            const char *f = strchr(pcs, c);
            if (f && c) {
                int idx = (int)(f - pcs), white = idx < 6, pc = idx % 6;
                int rank = 7 - row;
                if (rank < 0 || col > 7) return -1;
                u64 b = BIT(rank * 8 + col);
                if (white) out->white |= b;
                switch (pc) {
                case 0: out->pawns |= b; break;
                case 1: out->knights |= b; break;
                case 2: out->bishop_queens |= b; break;
                case 3: out->rook_queens |= b; break;
                case 4: out->bishop_queens |= b; out->rook_queens |= b; break;
                case 5: out->kings |= b; break;
                }
                col++;
                placed++;
            } else if (c == ' ') {
                return -1;
            }
        }
        if (row >= 7 && col == 8) break;
    }
    if (!fen[i]) return -1;
    i++;
    while (fen[i] == ' ') i++;
    set_chance(out, (fen[i] == 'b' || fen[i] == 'B') ? BLACK : WHITE);
    if (fen[i]) i++;
    while (fen[i] == ' ') i++;
    int wc = 0, bc = 0;
    while (fen[i] && fen[i] != ' ') {
        switch (fen[i]) {
        case 'k': bc |= CF_KING; break;
        case 'q': bc |= CF_QUEEN; break;
        case 'K': wc |= CF_KING; break;
        case 'Q': wc |= CF_QUEEN; break;
        }
        i++;
    }
    set_wc(out, wc);
    set_bc(out, bc);
    while (fen[i] == ' ') i++;
    if (fen[i] >= 'a' && fen[i] <= 'h') set_ep(out, fen[i] - 'a' + 1);
    if (__builtin_popcountll(out->kings & out->white) != 1 ||
        __builtin_popcountll(out->kings & ~out->white) != 1) return -1;
    (void)placed;
    return 0;
}

int orc_pos_to_fen(const orc_pos *p, char *buf, int buflen) {
    char tmp[128];
    int k = 0;
    u64 pawns = p->pawns & RANKS2TO7;
    for (int r = 7; r >= 0; r--) {
        int empty = 0;
        for (int f = 0; f < 8; f++) {
            u64 b = BIT(r * 8 + f);
            char c = 0;
            if (pawns & b) c = 'p';
            else if (p->knights & b) c = 'n';
            else if (p->kings & b) c = 'k';
            else if ((p->bishop_queens & b) && (p->rook_queens & b)) c = 'q';
            else if (p->bishop_queens & b) c = 'b';
            else if (p->rook_queens & b) c = 'r';
            if (!c) { empty++; continue; }
            if (empty) { tmp[k++] = (char)('0' + empty); empty = 0; }
            if (p->white & b) c = (char)(c - 32);
            tmp[k++] = c;
        }
        if (empty) tmp[k++] = (char)('0' + empty);
        if (r) tmp[k++] = '/';
    }
    tmp[k++] = ' ';
    tmp[k++] = st_chance(p) ? 'b' : 'w';
    tmp[k++] = ' ';
    int wc = st_wc(p), bc = st_bc(p);
    if (!wc && !bc) tmp[k++] = '-';
    if (wc & CF_KING) tmp[k++] = 'K';
    if (wc & CF_QUEEN) tmp[k++] = 'Q';
    if (bc & CF_KING) tmp[k++] = 'k';
    if (bc & CF_QUEEN) tmp[k++] = 'q';
    tmp[k++] = ' ';
    int ep = st_ep(p);
    if (ep) { tmp[k++] = (char)('a' + ep - 1); tmp[k++] = st_chance(p) ? '3' : '6'; }
    else tmp[k++] = '-';
    k += snprintf(tmp + k, sizeof(tmp) - (size_t)k, " 0 1");
    if (k + 1 > buflen) return -1;
    memcpy(buf, tmp, (size_t)k + 1);
    return k;
}

/* ---- Zobrist (build specification, see header) ---------------------------- */
void orc_zobrist_table(uint64_t seed, uint64_t *t) {
    u64 s = seed ? seed : 0x9E3779B97F4A7C15ULL;
    for (int i = 0; i < 781; i++) t[i] = xs64(&s);
}

/* key over the absolute position: pieces[color][type][sq] (type P,N,B,R,Q,K),
 * castle[color][king side, queen side] at 768.., ep[file] at 772.., side at
 * 780 XORed when white is to move (as MoveGeneratorBitboard.h:3849-3850). */
uint64_t orc_zobrist_key(const orc_pos *p, const uint64_t *t) {
    u64 k = 0;
    if (!st_chance(p)) k ^= t[780];
    int wc = st_wc(p), bc = st_bc(p), ep = st_ep(p);
    if (wc & CF_KING) k ^= t[768];
    if (wc & CF_QUEEN) k ^= t[769];
    if (bc & CF_KING) k ^= t[770];
    if (bc & CF_QUEEN) k ^= t[771];
    if (ep) k ^= t[772 + ep - 1];
    u64 pawns = p->pawns & RANKS2TO7;
    u64 all = pawns | p->knights | p->bishop_queens | p->rook_queens | p->kings;
    while (all) {
        int sq = __builtin_ctzll(all);
        u64 b = BIT(sq);
        int color = (p->white & b) ? 0 : 1, type;
        if (pawns & b) type = 0;
        else if (p->knights & b) type = 1;
        else if (p->kings & b) type = 5;
        else if ((p->bishop_queens & b) && (p->rook_queens & b)) type = 4;
        else if (p->bishop_queens & b) type = 2;
        else type = 3;
        k ^= t[(color * 6 + type) * 64 + sq];
        all &= all - 1;
    }
    return k;
}

/* ---- playouts (build specification) --------------------------------------- */
static int move_key(u16 m) {
    int f = m & 63, t = (m >> 6) & 63, fl = m >> 12, promo = 0;
    if (fl & 8) {
        switch (fl & 3) { case 0: promo = 1; break; case 1: promo = 2; break; case 2: promo = 3; break; default: promo = 4; }
    }
    return (f << 9) | (t << 3) | promo;
}

int orc_playouts(uint64_t seed, int n, int min_plies, int max_plies, int min_moves, int no_check, orc_pos *out) {
    orc_init();
    int o = g_backend; g_backend = ORC_KS;
    u64 s = seed ? seed : 0x9E3779B97F4A7C15ULL;
    orc_pos start;
    orc_fen_to_pos("rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1", &start);
    int got = 0;
    long attempts = 0;
    while (got < n) {
        if (++attempts > 1000000L + 1000L * n) break;
        int span = max_plies - min_plies + 1;
        int L = min_plies + (int)(xs64(&s) % (u64)span);
        orc_pos pos = start;
        int ok = 1;
        for (int ply = 0; ply < L; ply++) {
            u16 mv[256];
            u32 nm = orc_generate_moves(&pos, mv);
            if (nm == 0) { ok = 0; break; }
            /* canonical order: ascending (from, to, promotion piece) */
            for (u32 a = 1; a < nm; a++) {
                u16 x = mv[a]; int kx = move_key(x); int b = (int)a - 1;
                while (b >= 0 && move_key(mv[b]) > kx) { mv[b + 1] = mv[b]; b--; }
                mv[b + 1] = x;
            }
            u16 pick = mv[xs64(&s) % nm];
            orc_make_move(&pos, pick, 1);
        }
        if (!ok) continue;
        u32 nm = orc_count_moves(&pos);
        if (nm == 0) continue;
        if ((int)nm < min_moves) continue;
        if (no_check && orc_in_check(&pos)) continue;
        out[got++] = pos;
    }
    g_backend = o;
    return got;
}

/* ---- op model -------------------------------------------------------------- */
#ifdef ORC_OPCOUNT
static __thread u64 g_leaves __attribute__((tls_model("initial-exec")));
/* breadth-first path: make+count on every non-root node; generate on every
 * node with remaining depth >= 2 (the last interior ply). */
static void opcount_rec(const orc_pos *pos, int depth) {
    u16 mv[256];
    u32 n = orc_generate_moves(pos, mv);
    for (u32 i = 0; i < n; i++) {
        orc_pos c = *pos;
        orc_make_move(&c, mv[i], 0);
        u32 cnt = orc_count_moves(&c);
        if (depth == 2) g_leaves += cnt;
        else opcount_rec(&c, depth - 1);
    }
}
#endif

void orc_opcount_perft(const orc_pos *p, int depth, uint64_t *ops_out, uint64_t *leaves_out) {
#ifdef ORC_OPCOUNT
    int o = g_backend; g_backend = ORC_KS;
    g_ops = 0; g_leaves = 0;
    if (depth >= 2) {
        orc_count_moves(p);
        opcount_rec(p, depth);
    }
    *ops_out = g_ops; *leaves_out = g_leaves;
    g_backend = o;
#else
    (void)p; (void)depth;
    *ops_out = 0; *leaves_out = 0;
#endif
}
