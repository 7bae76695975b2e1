/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference perft path, used as the checker for the
 * HIP engine (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * are the only callers).  It is never linked into, loaded by, or called from
 * the product library.
 *
 * What it restates (reference file:line):
 *   position layout          chess.h:208-230        (HexaBitBoardPosition)
 *   CMove encoding           chess.h:234-293
 *   FEN parsing              util.cpp:189-280, util.cpp:92-140
 *   slider backends          MoveGeneratorBitboard.h:504-829 (KS, plain,
 *                            fixed-shift fancy, byte-lookup fancy)
 *   tables                   MoveGeneratorBitboard.h:982-1080, Magics.cpp:42-188
 *   pins / threats           MoveGeneratorBitboard.h:1213-1298
 *   countMoves               MoveGeneratorBitboard.h:2828-3295
 *   generateBoards           MoveGeneratorBitboard.h:1302-2261
 *   generateMoves            MoveGeneratorBitboard.h:2269-2823
 *   makeMove                 MoveGeneratorBitboard.h:3310-3560
 *   perft_bb (CPU perft)     perft_bb.h:79-101
 *
 * Defined here, not in the reference (the build's own specifications that the
 * product must match bit for bit): the seeded xorshift64 Zobrist tables, the
 * seeded random-playout position generator and the algorithmic op model.
 *
 * Parity pins (see tests/test_oracle.py): CPW known answers, an independent
 * mailbox generator (naive.c), brute-force slider tables, closed forms.
 */
#ifndef PERFT_ORACLE_H
#define PERFT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same 48-byte layout as the reference HexaBitBoardPosition (chess.h:208-230).
 * pawns bits 0-1 white castle, 2-3 black castle, 4-7 en-passant file+1,
 * 56-62 half-move counter, 63 side to move (1 = black). */
typedef struct orc_pos {
    uint64_t white;
    uint64_t pawns;
    uint64_t knights;
    uint64_t bishop_queens;
    uint64_t rook_queens;
    uint64_t kings;
} orc_pos;

enum { ORC_KS = 0, ORC_PLAIN = 1, ORC_FANCY = 2, ORC_BYTE = 3 };

int      orc_init(void);                      /* idempotent; 0 on success */
int      orc_set_backend(int backend);        /* returns previous backend */
int      orc_fen_to_pos(const char *fen, orc_pos *out);   /* 0 ok, -1 bad FEN */
int      orc_pos_to_fen(const orc_pos *p, char *buf, int buflen);

uint32_t orc_count_moves(const orc_pos *p);
uint32_t orc_generate_boards(const orc_pos *p, orc_pos *out);   /* out[256] */
uint32_t orc_generate_moves(const orc_pos *p, uint16_t *out);   /* out[256] */
void     orc_make_move(orc_pos *p, uint16_t move, int exact_ep);
int      orc_in_check(const orc_pos *p);

uint64_t orc_perft(const orc_pos *p, int depth);                 /* perft_bb */
/* perft of n positions, split over nthreads pthreads by position */
void     orc_perft_batch(const orc_pos *p, int n, int depth, int nthreads, uint64_t *out);
/* per-root-move divide: moves[i] (CMove) and counts[i]; returns #moves */
int      orc_divide(const orc_pos *p, int depth, uint16_t *moves, uint64_t *counts);

/* slider attacks of ONE piece on square sq given occupancy occ */
uint64_t orc_rook_attacks(int backend, int sq, uint64_t occ);
uint64_t orc_bishop_attacks(int backend, int sq, uint64_t occ);
uint64_t orc_rook_mask(int sq);
uint64_t orc_bishop_mask(int sq);
void     orc_table_stats(int *fancy_entries, int *rook_unique, int *bishop_unique);

/* Zobrist (build specification): xorshift64(13,7,17) stream seeded with
 * `seed` draws pieces[2][6][64], castle[2][2], ep[8], side -- 781 keys. */
uint64_t orc_xorshift64(uint64_t *state);
void     orc_zobrist_table(uint64_t seed, uint64_t *out781);
uint64_t orc_zobrist_key(const orc_pos *p, const uint64_t *table781);

/* Seeded random-playout positions (build specification). */
int      orc_playouts(uint64_t seed, int n, int min_plies, int max_plies,
                      int min_moves, int no_check, orc_pos *out);

/* Algorithmic op model (KS backend, int32 lane-op weights): ops of the
 * breadth-first path (generate at interior nodes, make+count at every
 * non-root node) summed over the whole tree of depth `depth`. */
void     orc_opcount_perft(const orc_pos *p, int depth, uint64_t *ops_out, uint64_t *leaves_out);

/* Independent mailbox legal-move generator (naive.c). */
uint64_t orc_naive_perft(const orc_pos *p, int depth);

#ifdef __cplusplus
}
#endif
#endif
