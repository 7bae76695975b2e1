/*
 * perft_gpu_amd.h -- C ABI of the MI355X perft engine.
 *
 * Plain pointers and sizes only; no C++ or torch types cross this boundary.
 * Each entry point names the reference interface it replaces (file:line in
 * sarona-RD/perft_gpu).
 *
 * Conventions (the reference printf()s and exit()s on errors and signals OOM
 * in-band with perfts[0] == ~0ull, perft_bb.h:1352-1363; here):
 *   - every call returns a pf_status; the library never prints or exits;
 *   - host arrays are caller-owned; the device arena, tables and the
 *     transposition table are library-owned;
 *   - one pf_ctx per GPU; calls on one context are serialised internally,
 *     pf_submit / pf_poll may be called from any thread while the serving loop
 *     runs.
 */
#ifndef PERFT_GPU_AMD_H
#define PERFT_GPU_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Position, byte-for-byte the reference HexaBitBoardPosition (chess.h:208-230):
 * six bitboards (a1 = bit 0); pawns bits 0-1 white castle (1 = O-O, 2 = O-O-O),
 * 2-3 black castle, 4-7 en-passant file + 1, 56-62 half-move counter (ignored),
 * 63 side to move (1 = black). */
typedef struct pf_pos48 {
    uint64_t white_pieces;
    uint64_t pawns;
    uint64_t knights;
    uint64_t bishop_queens;
    uint64_t rook_queens;
    uint64_t kings;
} pf_pos48;

typedef enum pf_status {
    PF_OK = 0,
    PF_EINVAL = -1,   /* bad argument, illegal FEN, depth out of range */
    PF_ENOMEM = -2,   /* device memory exhausted (reference: ALLSET sentinel) */
    PF_EDEVICE = -3,  /* HIP error or no usable GPU */
    PF_EBUSY = -4,    /* serving queue full, or the context is serving */
    PF_EAGAIN = -5,   /* nothing to report yet */
    PF_EUNAVAIL = -6, /* PF_MODE_PERSISTENT: the device refused the cooperative launch
                         (its grid-wide barrier resource is in use, e.g. by another
                         process); PF_MODE_AUTO falls back to the level driver */
    PF_EIO = -7       /* a file could not be opened, read or written */
} pf_status;

typedef enum pf_backend {
    PF_BACKEND_KS = 0,     /* Kogge-Stone sliders (USE_SLIDING_LUT 0) */
    PF_BACKEND_PLAIN = 1,  /* plain magics, 2.25 MiB (USE_FANCY_MAGICS 0) */
    PF_BACKEND_FANCY = 2,  /* fixed-shift fancy magics, 760 KiB (reference default) */
    PF_BACKEND_BYTE = 3    /* byte-lookup fancy magics, 146 KiB */
} pf_backend;

typedef enum pf_mode {
    PF_MODE_AUTO = 0,        /* TT configured: levels with the table; else the stream-ordered
                                single pass (PF_MODE_STREAM) when it fits, else levels */
    PF_MODE_LEVELS = 1,      /* host-driven level loop with depth-first chunking */
    PF_MODE_PERSISTENT = 2,  /* one cooperative launch for the whole batch (no TT) */
    PF_MODE_STREAM = 3       /* one launch per level queued on the stream, every level sized and
                                placed on the device: no host round trip, no grid barrier (no TT);
                                PF_ENOMEM when the batch does not fit the arena unchunked */
} pf_mode;

typedef struct pf_config {
    int backend;              /* pf_backend */
    int device;               /* HIP device ordinal */
    uint64_t arena_bytes;     /* frontier arena; 0 = 60% of free HBM after the TT */
    uint64_t tt_bytes;        /* transposition table; 0 = off */
    uint64_t zobrist_seed;    /* seed of the xorshift64 Zobrist tables; 0 = default */
    const char *tune_file;    /* autotuner result file to load; NULL = built-in */
} pf_config;

typedef struct pf_ctx pf_ctx;

typedef struct pf_result {
    uint64_t tag;        /* caller's tag from pf_submit */
    uint64_t nodes;      /* perft count */
    int32_t status;      /* pf_status */
    int32_t depth;
    double submit_s;     /* arrival time (library clock, seconds) */
    double done_s;       /* completion time */
} pf_result;

typedef struct pf_stats {
    uint64_t kernel_launches;
    uint64_t levels_run;
    uint64_t chunks;          /* depth-first chunk splits taken */
    uint64_t peak_arena_bytes;
    uint64_t tt_probes, tt_hits, tt_stores;  /* while serving: probes and hits since pf_serve_start */
    double gpu_ms;            /* device time of expand kernels (HIP events) */
    double last_expand_ms;    /* average duration of the last run's expand launches */
    uint64_t last_expand_launches;
    uint64_t dup_nodes;       /* nodes found to duplicate another node of their level (TT runs) */
    uint64_t frontier_nodes;  /* nodes written to frontier levels (TT runs) */
    uint64_t made_nodes;      /* children made and counted by the level kernels (TT runs): hits,
                                 duplicates, written nodes and the leaf pass's children */
    uint64_t graph_passes;    /* TT runs: chunks that ran as one captured-graph pass */
} pf_stats;

/* Replaces initGPU + MoveGeneratorBitboard::init + setupHashTables128b
 * (launcher.h:1506-1533, MoveGeneratorBitboard.h:982-1203, launcher.h:293-320). */
pf_status pf_init(const pf_config *cfg, pf_ctx **out);
/* Replaces MoveGeneratorBitboard::destroy / freeCompleteTT (MoveGeneratorBitboard.h:1205). */
void pf_destroy(pf_ctx *ctx);
const char *pf_strerror(int status);

/* Replaces Utils::readFENString + Utils::board088ToHexBB (util.cpp:189-280, 92-140):
 * same quirks -- the en-passant field keeps only its file, clocks are ignored. */
pf_status pf_fen_to_pos(const char *fen, pf_pos48 *out);
pf_status pf_pos_to_fen(const pf_pos48 *pos, char *buf, int buflen);

/* perft(position, depth): replaces perftLauncher (launcher.h:1455-1504), which
 * prints instead of returning.  depth 0 -> 1.  Plain per-position path: a
 * host-driven level loop, no cross-position batching. */
pf_status pf_perft(pf_ctx *ctx, const pf_pos48 *pos, int depth, uint64_t *out);

/* Batched perft of n positions (any mix of sides to move): replaces
 * perft_bb_gpu_simple_hash(count, positions, ..., perfts, depth, ...)
 * (perft_bb.h:1242), which needs one side to move and <= 256 roots.
 * `pos` and `out` may each be host memory or device memory of the context's
 * device (as in the reference, whose kernel takes device arrays).  With both
 * on the device, a table-free run (PF_MODE_STREAM, PF_MODE_PERSISTENT, or AUTO
 * with no table)
 * converts the roots on the device and leaves the counts there: no host
 * staging.  Other combinations stage through host memory.  A device pointer
 * of another device is PF_EINVAL. */
pf_status pf_perft_batch(pf_ctx *ctx, const pf_pos48 *pos, int n, int depth, uint64_t *out);
pf_status pf_perft_batch_mode(pf_ctx *ctx, const pf_pos48 *pos, int n, int depth, int mode, uint64_t *out);
/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The stream-ordered batch (PF_MODE_STREAM) with `pos` and `out`
 * both in device memory of the context's device (else PF_EINVAL); its control
 * block, converted roots and frontier levels live in one fixed 512 MiB device
 * array that the library holds for each device it runs on until the process
 * ends, resolved on the first call on that device, so the call allocates
 * nothing.  The counts accumulate in `out`.  PF_ENOMEM when the batch's
 * frontier exceeds the array (256 positions at perft 5 fit).  Every call on
 * one GPU uses the same array: calls that overlap on one GPU race. */
pf_status pf_perft_batch_noalloc(pf_ctx *ctx, const pf_pos48 *pos, int n, int depth, uint64_t *out);

/* Divided perft for the UCI "go perft": per legal root move, the move
 * (from | to << 6 | flag << 12, reference CMove flags without the capture
 * bit) and its perft(depth - 1).  moves/counts hold >= 256 entries. */
pf_status pf_divide(pf_ctx *ctx, const pf_pos48 *pos, int depth, uint16_t *moves, uint64_t *counts, int *n);
/* UCI text of a move in that encoding, e.g. "e7e8q"; buf holds >= 6 bytes. */
pf_status pf_move_to_uci(uint16_t move, char *buf);
/* Apply a UCI move text to a position (for "position ... moves ..."). */
pf_status pf_apply_uci(pf_pos48 *pos, const char *uci);

/* Slider backend selection at run time (the reference's switches.h:102-125
 * are compile-time). */
pf_status pf_set_backend(pf_ctx *ctx, int backend);
/* In-level duplicate elimination on runs with a transposition table (on by
 * default): the compile-time FIND_DUPLICATES_IN_BFS of switches.h:50-52 as a
 * run-time switch.  Counts are the same either way. */
pf_status pf_set_dedup(pf_ctx *ctx, int on);
/* Runs with a transposition table: a chunk of a level whose whole subtree is
 * predicted to fit the free frontier memory runs as one replay of a captured
 * hipGraph (on by default), every level below it sized and placed on the
 * device, with no host round trip per level (the reference syncs the device
 * after every level launch, perft_bb.h:1071-1135); off = the host-driven
 * level loop everywhere.  Counts are the same either way. */
pf_status pf_set_level_graphs(pf_ctx *ctx, int on);
/* En-passant flagging of the children made on runs with a table:
 * exact (on, the default) sets a child's en-passant file only when an enemy
 * pawn can capture (EXACT_EN_PASSENT_FLAGGING, switches.h:43-48,
 * MoveGeneratorBitboard.h:3528-3547); off sets it after every double push, as
 * the reference's CPU generator does (MoveGeneratorBitboard.h:1575-1582).
 * Counts are the same either way; keys, and so table hits and in-level
 * duplicates, are not.  PF_EBUSY while serving. */
pf_status pf_set_exact_ep(pf_ctx *ctx, int on);
/* Table layout: 0 (the default) = one table for every remaining depth (the
 * depth is verified in each entry); k = 1..4 = 2^k equal regions, an entry
 * going to the region of its remaining depth modulo 2^k (the reference's
 * one table per depth, launcher.h:62-126).  Empties the table.  PF_EINVAL
 * without a table or when a region would hold fewer than 1024 buckets;
 * pf_tt_save / pf_tt_load need layout 0; PF_EBUSY while serving. */
pf_status pf_set_tt_regions(pf_ctx *ctx, int log2_regions);
/* Autotuner (replaces the reference's compile-time launch shape, perft_bb.h:327-329,
 * switches.h:14-21, and its launch-depth estimate, launcher.h:1535-1563): times
 * the current backend on `sample` (n positions, `depth`) over
 *   - the table-free single passes: the cooperative kernel's tile shapes and
 *     workgroups per CU, the stream-ordered level kernels' forms (tile shape x
 *     waves per SIMD, leaf and write level separately; for the byte-lookup
 *     backend also the form with its tables staged in LDS, form 6) and which
 *     of the two passes PF_MODE_AUTO runs;
 *   - with a table: the level driver's write- and leaf-level workgroups per CU
 *     and, if the sample makes it chunk, the share of free memory a chunk's
 *     next level may take;
 * keeps the fastest and, if `path` is not NULL, writes every backend's choice
 * there (the file pf_config.tune_file loads).
 * Counts are checked identical across all tried shapes (PF_EDEVICE if not).
 * With a table configured, pf_tune empties it before every timed run and
 * leaves it empty: a table restored by pf_tt_load or filled by earlier runs is
 * discarded, so tune before loading one. */
pf_status pf_tune(pf_ctx *ctx, const pf_pos48 *sample, int n, int depth, const char *path);
/* The same, restricted to one part: PF_TUNE_SINGLE_PASS, PF_TUNE_TABLE, or
 * PF_TUNE_ALL (pf_tune).  Tuning the table path alone lets a large table
 * sample be timed without the table-free passes running on it too. */
typedef enum pf_tune_what { PF_TUNE_ALL = 0, PF_TUNE_SINGLE_PASS = 1, PF_TUNE_TABLE = 2 } pf_tune_what;
pf_status pf_tune_mode(pf_ctx *ctx, const pf_pos48 *sample, int n, int depth, int what, const char *path);

/* Records mode (replaces processPerftRecords, launcher.h:1578-1695): perft(depth)
 * of every record "<FEN ...> <occurrences>" of the text file `in_path`, appended
 * to "<in_path>.op" as "<record> <perft> <perft * occurrences>" in input order,
 * `batch` records per pf_perft_batch call (0 = 256).  Records that already
 * have an output line are skipped, so an interrupted run resumes; a torn last
 * output line is dropped first.  *weighted_total = the sum of the last column
 * over the whole output file; *records_done = records evaluated by this call.
 * PF_EINVAL for an unreadable record or an output file of another input. */
pf_status pf_records(pf_ctx *ctx, const char *in_path, int depth, int batch, uint64_t *weighted_total,
                     uint64_t *records_done);

/* Persistent transposition table (replaces the reference's disk hash,
 * diskTTProbe / diskTTStore and perfthash.dat, launcher.h:700-789, which keeps
 * deep subtree counts across runs).  pf_tt_save writes the context's whole
 * table to `path` (a 64-byte header, then the buckets as they lie in HBM).
 * pf_tt_load replaces the table's contents with a saved one: a table of the
 * same size is restored exactly; a larger one folds into this one keeping
 * each bucket's two deepest entries, a smaller one is replicated to every
 * bucket its keys can map to.  Entries stay verified by their second key, so
 * a restored table changes no count.  PF_EINVAL: no table in this context, not
 * a table file, a truncated file, or a table made with other Zobrist keys
 * (pf_config.zobrist_seed); PF_EIO: the file cannot be opened, read or
 * written; PF_EBUSY while serving. */
pf_status pf_tt_save(pf_ctx *ctx, const char *path);
pf_status pf_tt_load(pf_ctx *ctx, const char *path);

pf_status pf_get_stats(pf_ctx *ctx, pf_stats *out);
/* Wait for all device work of this context (the reference's cudaDeviceSynchronize
 * calls around its launches, launcher.h:1490-1492). */
pf_status pf_sync(pf_ctx *ctx);
pf_status pf_reset_stats(pf_ctx *ctx);

/* Serving (no reference counterpart; the reference answers one perft per
 * run): a background loop admits submitted requests while one of max_jobs
 * slots -- each an equal share of the frontier arena -- is free, advances all
 * admitted requests together (one captured-graph step per breadth-first
 * level), shares the transposition table between them and reports each
 * result with its arrival and completion times.  While serving, pf_perft*
 * and pf_divide on the same context return PF_EBUSY. */
pf_status pf_serve_start(pf_ctx *ctx, int max_jobs, uint64_t job_bytes);
pf_status pf_submit(pf_ctx *ctx, const pf_pos48 *pos, int depth, uint64_t tag);
/* Copies up to max finished results; returns how many (>= 0) or a pf_status. */
int pf_poll(pf_ctx *ctx, pf_result *out, int max);
/* Finishes every admitted and queued request, then stops the loop. */
pf_status pf_serve_stop(pf_ctx *ctx);
/* Seconds on the library clock used for pf_result times. */
double pf_now(void);

/* Diagnostics (no reference counterpart): measure a ceiling on this device.
 * kind 0: int32 VALU lane-ops/s; 1: streaming copy, read+write bytes/s over a
 * `bytes` buffer; 2: random 16-byte read-modify-writes/s over a `bytes` table;
 * 3: issue rate of one VALU opcode (lane-instructions/s over 8 independent
 * chains), `iters` selecting it: 0 v_add_u32, 1 v_xor_b32, 2 v_and_b32,
 * 3 v_bcnt_u32_b32, 4 v_lshlrev_b32, 5 v_mul_lo_u32, 6 v_lshlrev_b64,
 * 7 v_bitop3_b32, 8 v_and_or_b32, 9 v_or3_b32, 10 v_cndmask_b32,
 * 11 v_perm_b32, 12 v_add3_u32, 13 v_min_u32, 14 v_ashrrev_i32, 15 v_bfe_i32. */
pf_status pf_probe(pf_ctx *ctx, int kind, uint64_t bytes, uint64_t iters, double *out);

/* Seeded random-playout positions (workload generator, no reference
 * counterpart): xorshift64(13,7,17) stream; per candidate, a length uniform in
 * [min_plies, max_plies], each ply a uniform pick from the legal moves in
 * ascending (from, to, promotion N<B<R<Q) order; candidates ending in mate or
 * stalemate, with fewer than min_moves legal moves, or (no_check) in check are
 * discarded.  Returns PF_OK when n positions were produced. */
pf_status pf_playouts(uint64_t seed, int n, int min_plies, int max_plies, int min_moves, int no_check,
                      pf_pos48 *out);
/* Number of legal moves of a position (host generator; for tests and tools). */
int pf_count_moves_host(const pf_pos48 *pos);

#ifdef __cplusplus
}
#endif
#endif
