// convert.h -- conversion between the reference position layout (pf_pos48,
// absolute colours) and the engine's side-to-move Board.
#pragma once
#include "../../include/perft_gpu_amd.h"
#include "board.h"

namespace pf {

PF_HD int pos_black_to_move(const pf_pos48& p) { return (int)(p.pawns >> 63); }

PF_HD Board to_board(const pf_pos48& p) {
    const u64 pawns = p.pawns & kRanks2to7;
    const u64 occ = pawns | p.knights | p.bishop_queens | p.rook_queens | p.kings;
    const int wc = (int)(p.pawns & 3), bc = (int)((p.pawns >> 2) & 3), ep = (int)((p.pawns >> 4) & 15);
    Board b;
    if (!pos_black_to_move(p)) {
        b.us = p.white_pieces & occ;
        b.pawns = pawns | (u64)(wc | (bc << 2) | (ep << 4));
        b.knights = p.knights;
        b.bishops = p.bishop_queens;
        b.rooks = p.rook_queens;
        b.kings = p.kings;
    } else {
        b.us = flip(occ & ~p.white_pieces);
        b.pawns = flip(pawns) | (u64)(bc | (wc << 2) | (ep << 4));
        b.knights = flip(p.knights);
        b.bishops = flip(p.bishop_queens);
        b.rooks = flip(p.rook_queens);
        b.kings = flip(p.kings);
    }
    return b;
}

PF_HD pf_pos48 to_pos48(const Board& b, int black) {
    const u64 pawns = b.pawns & kRanks2to7;
    const u64 occ = pawns | b.knights | b.bishops | b.rooks | b.kings;
    const int r = (int)(b.pawns & 15), ep = (int)((b.pawns >> 4) & 15);
    pf_pos48 p;
    if (!black) {
        p.white_pieces = b.us;
        p.pawns = pawns | (u64)(r | (ep << 4));
        p.knights = b.knights;
        p.bishop_queens = b.bishops;
        p.rook_queens = b.rooks;
        p.kings = b.kings;
    } else {
        p.white_pieces = flip(occ & ~b.us);
        p.pawns = flip(pawns) | (u64)((r >> 2) | ((r & 3) << 2) | (ep << 4)) | (1ULL << 63);
        p.knights = flip(b.knights);
        p.bishop_queens = flip(b.bishops);
        p.rook_queens = flip(b.rooks);
        p.kings = flip(b.kings);
    }
    return p;
}

}  // namespace pf
