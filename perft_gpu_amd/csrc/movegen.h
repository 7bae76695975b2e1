// movegen.h -- legal move counting, generation and make-move on the
// side-to-move-is-us Board, for host and device.
//
// Semantics follow the reference move generator exactly, including its
// conventions on unreachable positions (MoveGeneratorBitboard.h):
//   * pins: an enemy slider on an empty-board king ray with exactly one piece
//     between pins that piece (findPinnedPieces, :1213-1249);
//   * threatened squares are computed with our king removed (x-ray,
//     findAttackedSquares, :1254-1298);
//   * en passant: a pinned capturer may capture only along its pin line; an
//     unpinned one is rejected only by the horizontal discovered-check test
//     (:3071-3106); in check, en passant counts only if the captured pawn is
//     the single checker (:2873-2890);
//   * in check, pinned pieces never move (:2867) and a pawn double push is
//     tried only if the single push does not resolve the check (:2906-2920);
//   * castling needs its right, empty F1G1 / B1D1 and unattacked F1G1 / C1D1
//     (:3180-3214);
//   * promotions count four (:3160-3178).
// The counting itself is organised for wave64 SIMT: pawns, knights and (with
// Kogge-Stone) sliders are counted set-wise with no per-piece loops.
#pragma once
#include "sliders.h"

namespace pf {

// Move encoding (16 bits): from | to << 6 | flag << 12.  Flags follow the
// reference CMove flags without the capture bit (chess.h:259-293).
enum : int {
    kFlQuiet = 0, kFlDouble = 1, kFlOO = 2, kFlOOO = 3, kFlEP = 5,
    kFlPromoN = 8, kFlPromoB = 9, kFlPromoR = 10, kFlPromoQ = 11
};
PF_HD u32 mk_move(int from, int to, int flag) { return (u32)(from | (to << 6) | (flag << 12)); }

struct PosInfo {
    u64 occ, us, them, pawns, K;
    u64 att;       // squares attacked by the opponent (through our king)
    u64 pinned;    // our pinned pieces
    u64 checkers;  // opponent pieces giving check
    u64 block;     // squares between king and a single sliding checker
    int ksq;
};

PF_HD u64 between_on(u64 line, int a, int b) {  // squares of `line` strictly between a and b
    int lo = a < b ? a : b, hi = a < b ? b : a;
    return line & (bit(hi) - (bit(lo) << 1));
}

// Everything of PosInfo but the threat map.  `xray` (when not null) receives,
// for each sliding checker, the square just beyond the king on its line: the
// one square that the threat map with our king removed (findAttackedSquares'
// x-ray) adds to a map computed with the king in place, among the squares a
// king move can reach.
PF_HD void pins_checks(const Board& b, PosInfo& pi, u64* xray) {
    const u64 pawns = b.pawns & kRanks2to7;
    const u64 occ = or3(or3(pawns, b.knights, b.bishops), b.rooks, b.kings);
    const u64 us = b.us, them = occ & ~us;
    const u64 K = b.kings & us;
    const int ksq = ctz(K);
    const u64 eB = b.bishops & them, eR = b.rooks & them;
    const u64 eP = pawns & them, eN = b.knights & them;
    u64 pinned = 0, checkers = 0, block = 0, beyond = 0;
    const u64 kf = file_of(ksq), kr = rank_of(ksq), kd = diag_of(ksq), ka = anti_of(ksq);
    const u64 kz = xray ? king_fill(K) : 0;
    u64 sn = either_and(kf, kr, eR) | either_and(kd, ka, eB);
    while (sn) {
        const int s = ctz(sn);
        sn &= sn - 1;
        const u64 sb = bit(s);
        const u64 line = (kf & sb) ? kf : (kr & sb) ? kr : (kd & sb) ? kd : ka;
        const u64 btw = between_on(line, ksq, s);
        const u64 in = btw & occ;
        if (!in) {
            checkers |= sb;
            block |= btw;
            beyond |= line & kz & ~(btw | sb);  // the king's neighbour on the line, away from the checker
        } else if (!(in & (in - 1))) {
            pinned |= in & us;
        }
    }
    checkers |= and_or(knight_fill(K), eN, (((K << 7) & kNotH) | ((K << 9) & kNotA)) & eP);
    if (xray) *xray = beyond;
    pi.occ = occ; pi.us = us; pi.them = them; pi.pawns = pawns; pi.K = K;
    pi.att = 0; pi.pinned = pinned; pi.checkers = checkers; pi.block = block; pi.ksq = ksq;
}

// Squares attacked by the opponent's pawns, knights and king.
PF_HD u64 leaper_threats(const Board& b, const PosInfo& pi) {
    const u64 eP = pi.pawns & pi.them;
    return or3(((eP >> 7) & kNotA) | ((eP >> 9) & kNotH), knight_fill(b.knights & pi.them), king_fill(b.kings & pi.them));
}

template <class SL>
PF_HD void analyse(const Board& b, const SL& sl, PosInfo& pi) {
    pins_checks(b, pi, nullptr);
    const u64 occX = pi.occ ^ pi.K;  // x-ray through our king (findAttackedSquares)
    pi.att = leaper_threats(b, pi) | sl.bishop_multi(b.bishops & pi.them, occX) | sl.rook_multi(b.rooks & pi.them, occX);
}

// Knight moves of every knight in `n` onto `tgt`, added to `acc`: the eight
// jump directions are translations, so two knights never share a target
// within a direction.  Each jump is one shift and one three-input and with
// its wrap mask.
PF_HD u32 knight_count(u64 n, u64 tgt, u32 acc) {
    constexpr u64 kNotAB = 0xfcfcfcfcfcfcfcfcULL, kNotGH = 0x3f3f3f3f3f3f3f3fULL;
    acc = popc_add(and3(shl<17>(n), kNotA, tgt), acc);
    acc = popc_add(and3(shl<15>(n), kNotH, tgt), acc);
    acc = popc_add(and3(shl<10>(n), kNotAB, tgt), acc);
    acc = popc_add(and3(shl<6>(n), kNotGH, tgt), acc);
    acc = popc_add(and3(shr<15>(n), kNotA, tgt), acc);
    acc = popc_add(and3(shr<17>(n), kNotH, tgt), acc);
    acc = popc_add(and3(shr<6>(n), kNotAB, tgt), acc);
    return popc_add(and3(shr<10>(n), kNotGH, tgt), acc);
}

// Slider moves onto `tgt` for bishops-like set B and rook-like set R, with
// pinned pieces kept on their axis.  Kogge-Stone: one fill per direction.
template <class SL>
PF_HD int slider_count(const SL& sl, u64 B, u64 R, u64 occ, u64 tgt, u64 pinned, int ksq) {
    if constexpr (SL::kSetwise) {
        const u64 empty = ~occ, free = ~pinned;
        const u64 rv = R & (free | file_of(ksq)), rh = R & (free | rank_of(ksq));
        const u64 bd = B & (free | diag_of(ksq)), ba = B & (free | anti_of(ksq));
        return popc(att_n(rv, empty) & tgt) + popc(att_s(rv, empty) & tgt) +
               popc(att_e(rh, empty) & tgt) + popc(att_w(rh, empty) & tgt) +
               popc(att_ne(bd, empty) & tgt) + popc(att_sw(bd, empty) & tgt) +
               popc(att_nw(ba, empty) & tgt) + popc(att_se(ba, empty) & tgt);
    } else {
        int n = 0;
        const u64 kl[4] = {file_of(ksq), rank_of(ksq), diag_of(ksq), anti_of(ksq)};
        while (B) {
            const int s = ctz(B);
            B &= B - 1;
            u64 a = sl.bishop(s, occ) & tgt;
            if (pinned & bit(s)) a &= (kl[2] & bit(s)) ? kl[2] : (kl[3] & bit(s)) ? kl[3] : 0;
            n += popc(a);
        }
        while (R) {
            const int s = ctz(R);
            R &= R - 1;
            u64 a = sl.rook(s, occ) & tgt;
            if (pinned & bit(s)) a &= (kl[0] & bit(s)) ? kl[0] : (kl[1] & bit(s)) ? kl[1] : 0;
            n += popc(a);
        }
        return n;
    }
}

// Number of legal moves (countMoves + countMovesOutOfCheck,
// MoveGeneratorBitboard.h:2828-3295).  One branch-free path serves both: out
// of check `safe` is every square; in single check it is the checker and
// the squares between it and the king, pinned pieces do not move and a
// double push counts only where the single push does not resolve the check;
// in double check `safe` is empty and only the king moves.  (A wave holds 64
// different positions, so a separate in-check path would run for nearly
// every wave.)
struct CountMasks {
    bool check;
    u64 safe;    // destinations that resolve a check (every square out of check)
    u64 tgt;     // non-king destinations: not ours, and safe
    u64 pinMov;  // pinned pieces that move along their pin line (none in check)
};
PF_HD CountMasks count_masks(const PosInfo& pi) {
    CountMasks m;
    m.check = pi.checkers != 0;
    const bool dbl = (pi.checkers & (pi.checkers - 1)) != 0;
    m.safe = !m.check ? ~0ULL : dbl ? 0ULL : (pi.checkers | pi.block);
    m.tgt = ~pi.us & m.safe;
    m.pinMov = m.check ? 0ULL : pi.pinned;
    return m;
}

// King moves and castling, given the threat map `att`, added to `n`.
PF_HD u32 count_king(const Board& b, const PosInfo& pi, const CountMasks& cm, u64 att, u32 n) {
    n = popc_add(andn2(king_fill(pi.K), pi.us, att), n);
    // castling (MoveGeneratorBitboard.h:3180-3198), never out of check
    const int r = st_rights(b);
    n += (!cm.check && (r & 1) && !(pi.occ & kF1G1) && !(att & kF1G1)) ? 1 : 0;
    n += (!cm.check && (r & 2) && !(pi.occ & kB1D1) && !(att & kC1D1)) ? 1 : 0;
    return n;
}

// Pawn, en-passant and knight moves, added to `n`.
PF_HD u32 count_pawns_knights(const Board& b, const PosInfo& pi, const CountMasks& cm, u32 n);

// Number of legal moves of the Kogge-Stone backend: the threat map's slider
// rays and our sliders' moves are filled direction by direction over one
// shared propagator chain, the map with our king in place plus the squares
// beyond it (pins_checks).
PF_HD u32 count_moves_ks(const Board& b) {
    PosInfo pi;
    u64 att = 0;
    pins_checks(b, pi, &att);
    const CountMasks cm = count_masks(pi);
    att |= leaper_threats(b, pi);
    const u64 them = pi.them, us = pi.us, empty = ~pi.occ, tgt = cm.tgt;
    const u64 eB = b.bishops & them, eR = b.rooks & them;
    // our sliders: pinned ones only along their pin line, none while in check
    const u64 frozen = pi.pinned & ~cm.pinMov, free = ~cm.pinMov;
    const u64 oB = and_andn(b.bishops, us, frozen), oR = and_andn(b.rooks, us, frozen);
    const int ksq = pi.ksq;
    const u64 rv = either_and(free, file_of(ksq), oR), rh = either_and(free, rank_of(ksq), oR);
    const u64 bd = either_and(free, diag_of(ksq), oB), ba = either_and(free, anti_of(ksq), oB);
    u32 n = 0;
    // each direction: the enemy ray (threat map) and our moves' count
#define PF_KS_DIR(D, C, E, O, RAY)                                              \
    {                                                                            \
        RAY = step<D>(ray_reach<D>(E, C));                                       \
        n = popc_add(and3(step<D>(ray_reach<D>(O, C)), tgt, step_mask<D>()), n); \
    }
    u64 rN, rS, rE, rW, rNE, rSW, rNW, rSE;
    {
        const RayChain s = ray_chain<-8>(empty), nn = north_chain(s);
        PF_KS_DIR(8, nn, eR, rv, rN)
        PF_KS_DIR(-8, s, eR, rv, rS)
    }
    {
        const RayChain c = ray_chain<1>(empty);
        PF_KS_DIR(1, c, eR, rh, rE)
    }
    {
        const RayChain c = ray_chain<-1>(empty);
        PF_KS_DIR(-1, c, eR, rh, rW)
    }
    {
        const RayChain c = ray_chain<9>(empty);
        PF_KS_DIR(9, c, eB, bd, rNE)
    }
    {
        const RayChain c = ray_chain<-9>(empty);
        PF_KS_DIR(-9, c, eB, bd, rSW)
    }
    {
        const RayChain c = ray_chain<7>(empty);
        PF_KS_DIR(7, c, eB, ba, rNW)
    }
    {
        const RayChain c = ray_chain<-7>(empty);
        PF_KS_DIR(-7, c, eB, ba, rSE)
    }
#undef PF_KS_DIR
    // the rays that still need the kNotA / kNotH wrap mask, then the map
    const u64 attA = or3(rE, rNE, rSE), attH = or3(rW, rSW, rNW);
    att = or_and(or_and(or3(att, rN, rS), attA, kNotA), attH, kNotH);
    return count_pawns_knights(b, pi, cm, count_king(b, pi, cm, att, n));
}

template <class SL>
PF_HD u32 count_moves(const Board& b, const SL& sl) {
    if constexpr (SL::kSetwise) {
        return count_moves_ks(b);
    } else {
        PosInfo pi;
        analyse(b, sl, pi);
        const CountMasks cm = count_masks(pi);
        const u64 frozen = pi.pinned & ~cm.pinMov;
        const u32 ns = (u32)slider_count(sl, b.bishops & pi.us & ~frozen, b.rooks & pi.us & ~frozen, pi.occ, cm.tgt,
                                         cm.pinMov, pi.ksq);
        return count_pawns_knights(b, pi, cm, count_king(b, pi, cm, pi.att, ns));
    }
}

PF_HD u32 count_pawns_knights(const Board& b, const PosInfo& pi, const CountMasks& cm, u32 n) {
    const u64 occ = pi.occ, us = pi.us, them = pi.them, empty = ~occ;
    const u64 P = pi.pawns & us;
    const u64 Pf = and_andn(pi.pawns, us, pi.pinned);
    const int ksq = pi.ksq;
    const bool check = cm.check;
    const u64 safe = cm.safe, pinMov = cm.pinMov;

    // pawns (MoveGeneratorBitboard.h:3109-3178; in check :2892-2940)
    {
        // Pinned pawns move only along their pin line: file-pinned ones push,
        // a1-h8-pinned ones capture up-right, a8-h1-pinned ones up-left.
        // pinMov is empty in check, and out of check `safe` is every square,
        // so they share the unpinned pawns' masks.
        const u64 pm = P & pinMov;
        const u64 push = or_and(Pf, pm, file_of(ksq));
        const u64 capR = or_and(Pf, pm, diag_of(ksq)), capL = or_and(Pf, pm, anti_of(ksq));
        // single and double pushes never share a target (a double push
        // needs the square in front of it empty, a single push to the same
        // square a pawn there), so they are counted in one popcount
        const u64 p1 = (push << 8) & empty;
        const u64 dmask = check ? ~safe : ~0ULL;
        const u64 p2 = (((p1 & dmask) & kRank3) << 8) & empty;
        n = popc_add(either_and(p1, p2, safe), n);
        const u64 ts = them & safe;
        const u64 cl = and3(shl<7>(capL), kNotH, ts), cr = and3(shl<9>(capR), kNotA, ts);
        n = popc_add(cr, popc_add(cl, n));
        // promotions count four: three more for each move onto the last rank,
        // the top byte of the board's upper half
        constexpr u32 kR8 = 0xFF000000u;
        u32 promo = popc_add32((u32)(cr >> 32) & kR8, 0u);
        promo = popc_add32((u32)(cl >> 32) & kR8, promo);
        promo = popc_add32((u32)((p1 & safe) >> 32) & kR8, promo);
        n += 3 * promo;
    }
    // en passant (MoveGeneratorBitboard.h:3058-3107; in check :2873-2890)
    const int ep = st_ep(b);
    if (ep) {
        const u64 T = bit(40 + ep - 1), C = T >> 8;
        const u64 src = (((C << 1) & kNotA) | ((C >> 1) & kNotH)) & P;
        if (check) {
            // only when the pawn that just moved is the single checker
            if (C == pi.checkers) n = popc_add(src & ~pi.pinned, n);
        } else {
            const u64 eR = b.rooks & them;
            u64 ok = src & ~pi.pinned;
            if ((pi.K & kRank5) && (eR & kRank5)) {
                // both pawns leave the king's rank: the horizontal discovered check test
                while (ok) {
                    const u64 s1 = ok & (0 - ok);
                    ok ^= s1;
                    const u64 pro = empty | C | s1;
                    n += ((att_e(eR, pro) | att_w(eR, pro)) & pi.K) ? 0 : 1;
                }
            } else {
                n = popc_add(ok, n);
            }
            u64 pin = src & pi.pinned;  // a pinned capturer only along its pin line
            while (pin) {
                const u64 s1 = pin & (0 - pin);
                pin ^= s1;
                const u64 pinF = pi.pinned & file_of(ksq), pinD = pi.pinned & diag_of(ksq);
                const u64 pinA = pi.pinned & anti_of(ksq);
                const u64 line = (pinD & s1) ? diag_of(ksq) : (pinA & s1) ? anti_of(ksq)
                                 : (pinF & s1) ? file_of(ksq) : rank_of(ksq);
                n += (T & line) ? 1 : 0;
            }
        }
    }
    // knights (pinned ones never move)
    return knight_count(and_andn(b.knights, us, pi.pinned), cm.tgt, n);
}

// Legal move generation.  `emit(move)` is called once per legal move; the
// order is: king, castles, knights, bishops, rooks, pawn pushes, double
// pushes, captures, en passant.  Perft counts do not depend on it.
template <class SL, class Emit>
PF_HD void generate(const Board& b, const SL& sl, const PosInfo& pi, Emit& emit) {
    const u64 occ = pi.occ, us = pi.us, them = pi.them, empty = ~occ;
    const int ksq = pi.ksq;
    const u64 P = pi.pawns & us, Pf = P & ~pi.pinned;
    {
        u64 t = king_fill(pi.K) & ~us & ~pi.att;
        while (t) { emit(mk_move(ksq, ctz(t), kFlQuiet)); t &= t - 1; }
    }
    const bool check = pi.checkers != 0;
    if (check && (pi.checkers & (pi.checkers - 1))) return;
    const u64 tgt = check ? (pi.checkers | pi.block) : ~us;
    const u64 free = ~pi.pinned;
    if (!check) {
        const int r = st_rights(b);
        if ((r & 1) && !(occ & kF1G1) && !(pi.att & kF1G1)) emit(mk_move(4, 6, kFlOO));
        if ((r & 2) && !(occ & kB1D1) && !(pi.att & kC1D1)) emit(mk_move(4, 2, kFlOOO));
    }
    {
        u64 N = b.knights & us & free;
        while (N) {
            const int s = ctz(N);
            N &= N - 1;
            u64 t = knight_fill(bit(s)) & tgt;
            while (t) { emit(mk_move(s, ctz(t), kFlQuiet)); t &= t - 1; }
        }
    }
    {
        const u64 kl[4] = {file_of(ksq), rank_of(ksq), diag_of(ksq), anti_of(ksq)};
        u64 S = (b.bishops | b.rooks) & us;
        if (check) S &= free;
        while (S) {
            const int s = ctz(S);
            const u64 sb = bit(s);
            S &= S - 1;
            u64 a = 0;
            if (b.bishops & sb) a |= sl.bishop(s, occ);
            if (b.rooks & sb) a |= sl.rook(s, occ);
            a &= tgt;
            if (pi.pinned & sb) a &= (kl[0] & sb) ? kl[0] : (kl[1] & sb) ? kl[1] : (kl[2] & sb) ? kl[2] : kl[3];
            while (a) { emit(mk_move(s, ctz(a), kFlQuiet)); a &= a - 1; }
        }
    }
    {
        const u64 pinF = pi.pinned & file_of(ksq), pinD = pi.pinned & diag_of(ksq), pinA = pi.pinned & anti_of(ksq);
        const u64 Ppush = check ? Pf : (Pf | (P & pinF));
        const u64 Pne = check ? Pf : (Pf | (P & pinD));
        const u64 Pnw = check ? Pf : (Pf | (P & pinA));
        const u64 p1 = (Ppush << 8) & empty;
        u64 t = p1 & tgt;
        while (t) {
            const int d = ctz(t);
            t &= t - 1;
            if (d >= 56) {
                emit(mk_move(d - 8, d, kFlPromoN)); emit(mk_move(d - 8, d, kFlPromoB));
                emit(mk_move(d - 8, d, kFlPromoR)); emit(mk_move(d - 8, d, kFlPromoQ));
            } else {
                emit(mk_move(d - 8, d, kFlQuiet));
            }
        }
        t = (((check ? (p1 & ~tgt) : p1) & kRank3) << 8) & empty & tgt;
        while (t) { const int d = ctz(t); t &= t - 1; emit(mk_move(d - 16, d, kFlDouble)); }
        for (int side = 0; side < 2; side++) {
            // a pawn pinned on a diagonal captures along it: its NE (a1-h8 pin) or
            // NW (a8-h1 pin) neighbour lies on the pin line by construction
            u64 c = side ? (((Pnw << 7) & kNotH) & them & tgt) : (((Pne << 9) & kNotA) & them & tgt);
            const int delta = side ? 7 : 9;
            while (c) {
                const int d = ctz(c);
                c &= c - 1;
                if (d >= 56) {
                    emit(mk_move(d - delta, d, kFlPromoN)); emit(mk_move(d - delta, d, kFlPromoB));
                    emit(mk_move(d - delta, d, kFlPromoR)); emit(mk_move(d - delta, d, kFlPromoQ));
                } else {
                    emit(mk_move(d - delta, d, kFlQuiet));
                }
            }
        }
        const int ep = st_ep(b);
        if (ep) {
            const u64 T = bit(40 + ep - 1), C = T >> 8;
            u64 src = (((C << 1) & kNotA) | ((C >> 1) & kNotH)) & P;
            const u64 eR = b.rooks & them;
            while (src) {
                const u64 s = src & (0 - src);
                src ^= s;
                bool ok;
                if (check) {
                    ok = (C == pi.checkers) && !(s & pi.pinned);
                } else if (s & pi.pinned) {
                    const u64 line = (pinD & s) ? diag_of(ksq) : (pinA & s) ? anti_of(ksq) : (pinF & s) ? file_of(ksq) : rank_of(ksq);
                    ok = (T & line) != 0;
                } else {
                    const u64 pro = empty | C | s;
                    ok = !((att_e(eR, pro) | att_w(eR, pro)) & pi.K);
                }
                if (ok) emit(mk_move(ctz(s), 40 + ep - 1, kFlEP));
            }
        }
    }
}

// Apply a legal move and return the child in ITS side-to-move frame
// (makeMove, MoveGeneratorBitboard.h:3310-3560, followed by the colour flip).
// Castling rights: any move from or to e1/h1/a1 (ours) or e8/h8/a8 (theirs)
// clears the matching rights -- the same rule as the reference's king-move,
// rook-move and rook-capture updates on every reachable position.  The child
// gets an en-passant file only if an enemy pawn stands next to the
// double-pushed pawn (EXACT_EN_PASSENT_FLAGGING, :3528-3547).
// The flag is decoded from its four bits, each widened to an all-ones or
// all-zero mask, so that no flag value costs a compare and selects:
//   bit 3: promotion, to N (bits 1-0 = 00), B (01), R (10) or Q (11);
//   bit 2: en passant (5 is the only flag with it);
//   bit 1 without bit 3: castling, short (2) or long (3);
//   bit 0 alone: double push (1).
PF_HD u32 flag_mask(u32 m, int k) { return (u32)((int)(m << (31 - k)) >> 31); }
PF_HD u64 widen(u32 x) { return ((u64)x << 32) | x; }

PF_HD Board make_move(const Board& b, u32 m) {
    const int from = m & 63, to = (m >> 6) & 63;
    const u32 f0 = flag_mask(m, 12), f1 = flag_mask(m, 13), f2 = flag_mask(m, 14), f3 = flag_mask(m, 15);
    const u64 src = bit(from), dst = bit(to), mv = src | dst;
    u64 P = and_andn(b.pawns, kRanks2to7, dst);
    u64 N = b.knights & ~dst, B = b.bishops & ~dst, R = b.rooks & ~dst, Kg = b.kings;
    u64 us = xor_or(b.us, src, dst);
    P = xor_and(P, mv, 0 - ((P >> from) & 1));
    N = xor_and(N, mv, 0 - ((N >> from) & 1));
    B = xor_and(B, mv, 0 - ((B >> from) & 1));
    R = xor_and(R, mv, 0 - ((R >> from) & 1));
    Kg = xor_and(Kg, mv, 0 - ((Kg >> from) & 1));
    // promotion: the pawn leaves, the new piece appears on dst
    P = bitop3<0x70>(P, dst, widen(f3));                                   // P & ~(dst & f3)
    N = or_and(N, dst, widen(bitop3_32<0x10>(f3, f1, f0)));               // f3 & ~f1 & ~f0
    B = or_and(B, dst, widen(f3 & f0));
    R = or_and(R, dst, widen(f3 & f1));
    // en passant: the captured pawn stands behind dst
    P = bitop3<0x70>(P, shr<8>(dst), widen(f2));
    // castling: the rook's two squares (h1, f1 or a1, d1) toggle in R and us
    const u32 rook = bitop3_32<0x40>(bitop3_32<0xCA>(f0, 0x09u, 0xA0u), f1, f3);  // (f0 ? 0x09 : 0xA0) & f1 & ~f3
    R ^= rook;
    us ^= rook;
    int rights = (int)(b.pawns & 15);
    const int clr = (int)(((mv >> 4) & 1) * 3) | (int)((mv >> 7) & 1) | (int)((mv & 1) << 1) |
                    (int)(((mv >> 60) & 1) * 12) | (int)(((mv >> 63) & 1) << 2) | (int)(((mv >> 56) & 1) << 3);
    rights &= ~clr;
    // double push: an en-passant file only if an enemy pawn stands beside the
    // pawn on its fourth rank (the low half of the board)
    const u32 d4 = (u32)dst, side = bitop3_32<0xF8>((d4 >> 1) & 0x7F7F7F7Fu, d4 << 1, 0xFEFEFEFEu);
    const u32 dbl = bitop3_32<0x10>(f0, f1, f2 | f3);                     // f0 & ~f1 & ~(f2 | f3)
    const int ep = (bitop3_32<0x40>(side, (u32)P, (u32)us) & dbl) ? (from & 7) + 1 : 0;
    const u64 occ = or3(or3(P, N, B), R, Kg);
    Board c;
    c.us = flip(occ & ~us);
    c.pawns = flip(P) | (u64)(((rights >> 2) & 3) | ((rights & 3) << 2)) | ((u64)ep << 4);
    c.knights = flip(N);
    c.bishops = flip(B);
    c.rooks = flip(R);
    c.kings = flip(Kg);
    return c;
}

PF_HD Board with_count(Board b, u32 n) {
    b.pawns = (b.pawns & ~(0xFFULL << 56)) | ((u64)n << 56);
    return b;
}

}  // namespace pf
