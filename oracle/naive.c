/*
 * naive.c -- TEST INFRASTRUCTURE ONLY.
 *
 * An independent legal-move counter used to pin the oracle: a square-by-square
 * mailbox board, pseudo-legal generation by ray walking, and legality by
 * "make the move, then ask whether the mover's king is attacked".  It shares
 * no code and no algorithm with oracle.c (no bitboards, no pins, no check
 * evasion logic), so a dropped term or a wrong shift in the bitboard path shows
 * up as a count mismatch on random positions.  Slow; for small depths only.
 */
#include "oracle.h"

#include <string.h>

typedef struct {
    signed char sq[64]; /* 0 empty; +1..+6 white P N B R Q K; -1..-6 black */
    int side;           /* 0 white, 1 black */
    int castle;         /* bit0 K, bit1 Q, bit2 k, bit3 q */
    int ep_file;        /* -1 none, else file of the pawn that just double-pushed */
} mbox;

static void from_pos(const orc_pos *p, mbox *b) {
    memset(b, 0, sizeof(*b));
    uint64_t pawns = p->pawns & 0x00FFFFFFFFFFFF00ULL;
    for (int s = 0; s < 64; s++) {
        uint64_t m = 1ULL << s;
        int t = 0;
        if (pawns & m) t = 1;
        else if (p->knights & m) t = 2;
        else if (p->kings & m) t = 6;
        else if ((p->bishop_queens & m) && (p->rook_queens & m)) t = 5;
        else if (p->bishop_queens & m) t = 3;
        else if (p->rook_queens & m) t = 4;
        if (t) b->sq[s] = (signed char)((p->white & m) ? t : -t);
    }
    b->side = (int)(p->pawns >> 63);
    int wc = (int)(p->pawns & 3), bc = (int)((p->pawns >> 2) & 3);
    b->castle = (wc & 1) | ((wc & 2)) | ((bc & 1) << 2) | ((bc & 2) << 2);
    int ep = (int)((p->pawns >> 4) & 15);
    b->ep_file = ep ? ep - 1 : -1;
}

static int on_board(int f, int r) { return f >= 0 && f < 8 && r >= 0 && r < 8; }

/* is square s attacked by side `by` (0 white, 1 black)? */
static int attacked(const mbox *b, int s, int by) {
    int f = s & 7, r = s >> 3, sg = by ? -1 : 1;
    /* pawns: a white pawn attacks up, so look one rank down from s */
    int pr = r - sg;
    for (int df = -1; df <= 1; df += 2)
        if (on_board(f + df, pr) && b->sq[pr * 8 + f + df] == sg * 1) return 1;
    static const int kn[8][2] = {{1, 2}, {2, 1}, {2, -1}, {1, -2}, {-1, -2}, {-2, -1}, {-2, 1}, {-1, 2}};
    for (int i = 0; i < 8; i++) {
        int ff = f + kn[i][0], rr = r + kn[i][1];
        if (on_board(ff, rr) && b->sq[rr * 8 + ff] == sg * 2) return 1;
    }
    for (int df = -1; df <= 1; df++)
        for (int dr = -1; dr <= 1; dr++) {
            if (!df && !dr) continue;
            int ff = f + df, rr = r + dr;
            if (on_board(ff, rr) && b->sq[rr * 8 + ff] == sg * 6) return 1;
            int diag = df && dr;
            ff = f + df; rr = r + dr;
            while (on_board(ff, rr)) {
                int pc = b->sq[rr * 8 + ff];
                if (pc) {
                    if (pc == sg * 5 || pc == sg * (diag ? 3 : 4)) return 1;
                    break;
                }
                ff += df; rr += dr;
            }
        }
    return 0;
}

static int king_sq(const mbox *b, int side) {
    int k = side ? -6 : 6;
    for (int s = 0; s < 64; s++) if (b->sq[s] == k) return s;
    return -1;
}

typedef struct { int from, to, promo, ep, castle; } nmove;

static int gen_pseudo(const mbox *b, nmove *out) {
    int n = 0, side = b->side, sg = side ? -1 : 1;
    for (int s = 0; s < 64; s++) {
        int pc = b->sq[s];
        if (!pc || (pc > 0) != (side == 0)) continue;
        int t = pc * sg, f = s & 7, r = s >> 3;
        if (t == 1) {
            int r1 = r + sg, last = side ? 0 : 7, start = side ? 6 : 1;
            if (on_board(f, r1) && !b->sq[r1 * 8 + f]) {
                if (r1 == last) for (int pr = 2; pr <= 5; pr++) out[n++] = (nmove){s, r1 * 8 + f, pr, 0, 0};
                else out[n++] = (nmove){s, r1 * 8 + f, 0, 0, 0};
                int r2 = r + 2 * sg;
                if (r == start && !b->sq[r2 * 8 + f]) out[n++] = (nmove){s, r2 * 8 + f, 0, 0, 0};
            }
            for (int df = -1; df <= 1; df += 2) {
                int ff = f + df;
                if (!on_board(ff, r1)) continue;
                int d = r1 * 8 + ff, q = b->sq[d];
                if (q && (q > 0) != (side == 0)) {
                    if (r1 == last) for (int pr = 2; pr <= 5; pr++) out[n++] = (nmove){s, d, pr, 0, 0};
                    else out[n++] = (nmove){s, d, 0, 0, 0};
                }
                /* en passant: target square behind the double-pushed pawn */
                if (b->ep_file == ff && r == (side ? 3 : 4) && b->sq[r * 8 + ff] == -sg && !q)
                    out[n++] = (nmove){s, d, 0, 1, 0};
            }
        } else if (t == 2 || t == 6) {
            static const int kn[8][2] = {{1, 2}, {2, 1}, {2, -1}, {1, -2}, {-1, -2}, {-2, -1}, {-2, 1}, {-1, 2}};
            static const int kg[8][2] = {{1, 0}, {1, 1}, {0, 1}, {-1, 1}, {-1, 0}, {-1, -1}, {0, -1}, {1, -1}};
            const int (*dd)[2] = t == 2 ? kn : kg;
            for (int i = 0; i < 8; i++) {
                int ff = f + dd[i][0], rr = r + dd[i][1];
                if (!on_board(ff, rr)) continue;
                int q = b->sq[rr * 8 + ff];
                if (q && (q > 0) == (side == 0)) continue;
                out[n++] = (nmove){s, rr * 8 + ff, 0, 0, 0};
            }
        } else {
            for (int df = -1; df <= 1; df++)
                for (int dr = -1; dr <= 1; dr++) {
                    if (!df && !dr) continue;
                    int diag = df && dr;
                    if (t == 3 && !diag) continue;
                    if (t == 4 && diag) continue;
                    int ff = f + df, rr = r + dr;
                    while (on_board(ff, rr)) {
                        int q = b->sq[rr * 8 + ff];
                        if (q && (q > 0) == (side == 0)) break;
                        out[n++] = (nmove){s, rr * 8 + ff, 0, 0, 0};
                        if (q) break;
                        ff += df; rr += dr;
                    }
                }
        }
    }
    /* castling: king and rook on their squares, path empty, king not in check
     * and not crossing or landing on an attacked square */
    int base = side ? 56 : 0, k = 6 * sg, rk = 4 * sg, opp = !side;
    int kbit = side ? 4 : 1, qbit = side ? 8 : 2;
    if (b->sq[base + 4] == k) {
        if ((b->castle & kbit) && b->sq[base + 7] == rk && !b->sq[base + 5] && !b->sq[base + 6] &&
            !attacked(b, base + 4, opp) && !attacked(b, base + 5, opp) && !attacked(b, base + 6, opp))
            out[n++] = (nmove){base + 4, base + 6, 0, 0, 1};
        if ((b->castle & qbit) && b->sq[base + 0] == rk && !b->sq[base + 1] && !b->sq[base + 2] && !b->sq[base + 3] &&
            !attacked(b, base + 4, opp) && !attacked(b, base + 3, opp) && !attacked(b, base + 2, opp))
            out[n++] = (nmove){base + 4, base + 2, 0, 0, 2};
    }
    return n;
}

static void make(const mbox *b, const nmove *m, mbox *o) {
    *o = *b;
    int sg = b->side ? -1 : 1, pc = o->sq[m->from];
    o->sq[m->from] = 0;
    if (m->ep) o->sq[(m->from & ~7) | (m->to & 7)] = 0;
    o->sq[m->to] = (signed char)(m->promo ? m->promo * sg : pc);
    if (m->castle == 1) { o->sq[m->to + 1] = 0; o->sq[m->to - 1] = (signed char)(4 * sg); }
    if (m->castle == 2) { o->sq[m->to - 2] = 0; o->sq[m->to + 1] = (signed char)(4 * sg); }
    /* rights: any move from or to a corner or the king square clears them */
    static const int corner_bit[64] = {[0] = 2, [7] = 1, [56] = 8, [63] = 4, [4] = 3, [60] = 12};
    o->castle &= ~(corner_bit[m->from] | corner_bit[m->to]);
    o->ep_file = -1;
    if (pc * sg == 1 && (m->to - m->from == 16 || m->from - m->to == 16)) o->ep_file = m->from & 7;
    o->side = !b->side;
}

static uint64_t rec(const mbox *b, int depth) {
    nmove mv[320];
    int n = gen_pseudo(b, mv);
    uint64_t c = 0;
    for (int i = 0; i < n; i++) {
        mbox o;
        make(b, &mv[i], &o);
        int ks = king_sq(&o, b->side);
        if (ks < 0 || attacked(&o, ks, o.side)) continue;
        c += depth <= 1 ? 1 : rec(&o, depth - 1);
    }
    return c;
}

uint64_t orc_naive_perft(const orc_pos *p, int depth) {
    if (depth <= 0) return 1;
    mbox b;
    from_pos(p, &b);
    return rec(&b, depth);
}
