// capi.cpp -- the extern "C" boundary (include/perft_gpu_amd.h).
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/perft_gpu_amd.h"
#include "engine.h"
#include "serve.h"

struct pf_ctx {
    pf::Engine eng;
    pf::Server* srv = nullptr;
    // guards srv for pf_submit / pf_poll, which run without the context lock
    // while pf_serve_stop / pf_destroy may delete it
    std::mutex srv_mu;
};

using namespace pf;

namespace {

// Host instance of the move generator (Kogge-Stone, no tables) for set-up
// work around the GPU path: FEN checks, root moves of a divide, playouts.
const KSBackend kHostSl{};

int abs_sq(int sq, int black) { return black ? sq ^ 56 : sq; }

u32 to_abs_move(u32 m, int black) {
    const int from = abs_sq(m & 63, black), to = abs_sq((m >> 6) & 63, black);
    return (u32)from | ((u32)to << 6) | (m & 0xF000u);
}

struct VecEmit {
    u32* v;
    int n;
    void operator()(u32 m) { v[n++] = m; }
};

int host_moves(const Board& b, u32* out) {
    PosInfo pi;
    analyse(b, kHostSl, pi);
    VecEmit e{out, 0};
    generate(b, kHostSl, pi, e);
    return e.n;
}

int promo_code(u32 m) {
    const int fl = (m >> 12) & 15;
    if (!(fl & 8)) return 0;
    return 1 + (fl & 3);  // N=1, B=2, R=3, Q=4
}

}  // namespace

extern "C" {

const char* pf_strerror(int s) {
    switch (s) {
    case PF_OK: return "ok";
    case PF_EINVAL: return "invalid argument";
    case PF_ENOMEM: return "out of device memory";
    case PF_EDEVICE: return "device error";
    case PF_EBUSY: return "busy";
    case PF_EAGAIN: return "no result yet";
    case PF_EUNAVAIL: return "cooperative launch unavailable";
    case PF_EIO: return "file input/output error";
    default: return "unknown status";
    }
}

pf_status pf_init(const pf_config* cfg, pf_ctx** out) {
    if (!out) return PF_EINVAL;
    *out = nullptr;
    pf_config c{};
    if (cfg) c = *cfg;
    pf_ctx* ctx = new (std::nothrow) pf_ctx;
    if (!ctx) return PF_ENOMEM;
    pf_status s = ctx->eng.init(c);
    if (s == PF_OK) s = ctx->eng.init_tt(c.tt_bytes, c.zobrist_seed);
    if (s == PF_OK && c.tune_file) s = ctx->eng.load_tuning(c.tune_file);
    if (s != PF_OK) {
        ctx->eng.destroy();
        delete ctx;
        return s;
    }
    *out = ctx;
    return PF_OK;
}

void pf_destroy(pf_ctx* ctx) {
    if (!ctx) return;
    if (ctx->srv) {
        ctx->srv->stop();
        pf::Server* s = nullptr;
        {
            std::lock_guard<std::mutex> g(ctx->srv_mu);
            std::swap(s, ctx->srv);
        }
        delete s;
    }
    ctx->eng.destroy();
    delete ctx;
}

// util.cpp:189-280 (readFENString) + util.cpp:92-140 (board088ToHexBB)
pf_status pf_fen_to_pos(const char* fen, pf_pos48* out) {
    if (!fen || !out) return PF_EINVAL;
    memset(out, 0, sizeof(*out));
    int i = 0, row = 0, col = 0;
    for (; fen[i]; i++) {
        const char c = fen[i];
        if (c == '/' || c == '\\') { row++; col = 0; }
        else if (c >= '1' && c <= '8') col += c - '0';
        else {
            const char* pcs = "PNBRQKpnbrqk";
            const char* f = c ? strchr(pcs, c) : nullptr;
            if (!f) return PF_EINVAL;
            const int idx = (int)(f - pcs), pc = idx % 6, rank = 7 - row;
            if (rank < 0 || col > 7) return PF_EINVAL;
            const u64 b = bit(rank * 8 + col);
            if (idx < 6) out->white_pieces |= b;
            if (pc == 0) out->pawns |= b;
            if (pc == 1) out->knights |= b;
            if (pc == 2 || pc == 4) out->bishop_queens |= b;
            if (pc == 3 || pc == 4) out->rook_queens |= b;
            if (pc == 5) out->kings |= b;
            col++;
        }
        if (row >= 7 && col == 8) break;
    }
    if (!fen[i]) return PF_EINVAL;
    i++;
    while (fen[i] == ' ') i++;
    const int black = (fen[i] == 'b' || fen[i] == 'B');
    if (fen[i]) i++;
    while (fen[i] == ' ') i++;
    int wc = 0, bc = 0;
    while (fen[i] && fen[i] != ' ') {
        if (fen[i] == 'K') wc |= 1;
        if (fen[i] == 'Q') wc |= 2;
        if (fen[i] == 'k') bc |= 1;
        if (fen[i] == 'q') bc |= 2;
        i++;
    }
    while (fen[i] == ' ') i++;
    int ep = 0;
    if (fen[i] >= 'a' && fen[i] <= 'h') ep = fen[i] - 'a' + 1;
    if (out->pawns & ~kRanks2to7) return PF_EINVAL;
    if (popc(out->kings & out->white_pieces) != 1 || popc(out->kings & ~out->white_pieces) != 1) return PF_EINVAL;
    out->pawns |= (u64)(wc | (bc << 2) | (ep << 4)) | ((u64)black << 63);
    return PF_OK;
}

pf_status pf_pos_to_fen(const pf_pos48* p, char* buf, int buflen) {
    if (!p || !buf || buflen < 90) return PF_EINVAL;
    int k = 0;
    const u64 pawns = p->pawns & kRanks2to7;
    for (int r = 7; r >= 0; r--) {
        int empty = 0;
        for (int f = 0; f < 8; f++) {
            const u64 b = bit(r * 8 + f);
            char c = 0;
            if (pawns & b) c = 'p';
            else if (p->knights & b) c = 'n';
            else if (p->kings & b) c = 'k';
            else if ((p->bishop_queens & b) && (p->rook_queens & b)) c = 'q';
            else if (p->bishop_queens & b) c = 'b';
            else if (p->rook_queens & b) c = 'r';
            if (!c) { empty++; continue; }
            if (empty) { buf[k++] = (char)('0' + empty); empty = 0; }
            buf[k++] = (p->white_pieces & b) ? (char)(c - 32) : c;
        }
        if (empty) buf[k++] = (char)('0' + empty);
        if (r) buf[k++] = '/';
    }
    const int black = (int)(p->pawns >> 63), wc = (int)(p->pawns & 3), bc = (int)((p->pawns >> 2) & 3);
    const int ep = (int)((p->pawns >> 4) & 15);
    buf[k++] = ' ';
    buf[k++] = black ? 'b' : 'w';
    buf[k++] = ' ';
    if (!wc && !bc) buf[k++] = '-';
    if (wc & 1) buf[k++] = 'K';
    if (wc & 2) buf[k++] = 'Q';
    if (bc & 1) buf[k++] = 'k';
    if (bc & 2) buf[k++] = 'q';
    buf[k++] = ' ';
    if (ep) { buf[k++] = (char)('a' + ep - 1); buf[k++] = black ? '3' : '6'; }
    else buf[k++] = '-';
    memcpy(buf + k, " 0 1", 5);
    return PF_OK;
}

static pf_status check_depth(int depth) { return (depth < 0 || depth > 16) ? PF_EINVAL : PF_OK; }

pf_status pf_perft(pf_ctx* ctx, const pf_pos48* pos, int depth, uint64_t* out) {
    if (!ctx || !pos || !out || check_depth(depth)) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    if (ctx->eng.serving) return PF_EBUSY;
    return ctx->eng.perft_levels(pos, 1, depth, out);
}

// Where a caller's buffer lives: 0 host (pageable, pinned or managed), 1 device
// memory of `device`, -1 device memory of another device.
static int buffer_kind(const void* p, int device) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory the runtime has never seen
        return 0;
    }
    if (a.type != hipMemoryTypeDevice) return 0;
    return a.device == device ? 1 : -1;
}

// The table-free single passes: PERSISTENT the cooperative launch, STREAM the
// stream-ordered level launches, AUTO the one the tuner chose (stream until
// tuned).
static pf_status single_pass(pf_ctx* ctx, const pf_pos48* pos, int n, int depth, int mode, uint64_t* out, bool din,
                             bool dout) {
    const bool coop = mode == PF_MODE_PERSISTENT ||
                      (mode == PF_MODE_AUTO && ctx->eng.tune[ctx->eng.backend].single_pass == 1);
    if (coop) return ctx->eng.perft_persistent(pos, n, depth, out, din, dout);
    return ctx->eng.perft_stream(pos, n, depth, out, din, dout);
}

static pf_status batch_host(pf_ctx* ctx, const pf_pos48* pos, int n, int depth, int mode, uint64_t* out) {
    // the transposition table lives in the level driver; AUTO uses it when
    // configured, else a single pass that falls back to the level driver
    if (mode == PF_MODE_LEVELS || (mode == PF_MODE_AUTO && ctx->eng.d_tt && depth >= 3))
        return ctx->eng.perft_levels(pos, n, depth, out);
    pf_status s = single_pass(ctx, pos, n, depth, mode, out, false, false);
    if ((s == PF_ENOMEM || s == PF_EUNAVAIL) && mode == PF_MODE_AUTO) s = ctx->eng.perft_levels(pos, n, depth, out);
    return s;
}

pf_status pf_perft_batch_mode(pf_ctx* ctx, const pf_pos48* pos, int n, int depth, int mode, uint64_t* out) {
    if (!ctx || (n > 0 && (!pos || !out)) || n < 0 || check_depth(depth)) return PF_EINVAL;
    if (mode != PF_MODE_AUTO && mode != PF_MODE_LEVELS && mode != PF_MODE_PERSISTENT && mode != PF_MODE_STREAM)
        return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    if (ctx->eng.serving) return PF_EBUSY;
    if (n == 0) return PF_OK;
    if (hipSetDevice(ctx->eng.device) != hipSuccess) return PF_EDEVICE;
    const int kin = buffer_kind(pos, ctx->eng.device), kout = buffer_kind(out, ctx->eng.device);
    if (kin < 0 || kout < 0) return PF_EINVAL;
    if (!kin && !kout) return batch_host(ctx, pos, n, depth, mode, out);
    const bool table_run = mode == PF_MODE_LEVELS || (mode == PF_MODE_AUTO && ctx->eng.d_tt && depth >= 3);
    if (kin && kout && !table_run && depth >= 2) {
        pf_status s = single_pass(ctx, pos, n, depth, mode, out, true, true);
        if (!((s == PF_ENOMEM || s == PF_EUNAVAIL) && mode == PF_MODE_AUTO)) return s;
    }
    // staged: device buffers round-trip through host memory around the host path
    std::vector<pf_pos48> hp;
    std::vector<uint64_t> ho;
    const pf_pos48* p = pos;
    uint64_t* o = out;
    if (kin) {
        hp.resize((size_t)n);
        if (hipMemcpy(hp.data(), pos, (size_t)n * sizeof(pf_pos48), hipMemcpyDeviceToHost) != hipSuccess)
            return PF_EDEVICE;
        p = hp.data();
    }
    if (kout) {
        ho.resize((size_t)n);
        o = ho.data();
    }
    pf_status s = batch_host(ctx, p, n, depth, mode, o);
    if (s == PF_OK && kout &&
        hipMemcpy(out, ho.data(), (size_t)n * sizeof(uint64_t), hipMemcpyHostToDevice) != hipSuccess)
        return PF_EDEVICE;
    return s;
}

pf_status pf_perft_batch(pf_ctx* ctx, const pf_pos48* pos, int n, int depth, uint64_t* out) {
    return pf_perft_batch_mode(ctx, pos, n, depth, PF_MODE_AUTO, out);
}

pf_status pf_perft_batch_noalloc(pf_ctx* ctx, const pf_pos48* pos, int n, int depth, uint64_t* out) {
    if (!ctx || (n > 0 && (!pos || !out)) || n < 0 || check_depth(depth)) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    if (ctx->eng.serving) return PF_EBUSY;
    if (n == 0) return PF_OK;
    if (hipSetDevice(ctx->eng.device) != hipSuccess) return PF_EDEVICE;
    // both buffers in device memory of this context's device: nothing is staged
    if (buffer_kind(pos, ctx->eng.device) != 1 || buffer_kind(out, ctx->eng.device) != 1) return PF_EINVAL;
    return ctx->eng.perft_noalloc(pos, n, depth, out);
}

pf_status pf_divide(pf_ctx* ctx, const pf_pos48* pos, int depth, uint16_t* moves, uint64_t* counts, int* n) {
    if (!ctx || !pos || !moves || !counts || !n || depth < 1 || check_depth(depth)) return PF_EINVAL;
    const int black = pos_black_to_move(*pos);
    const Board b = to_board(*pos);
    u32 mv[256];
    const int nm = host_moves(b, mv);
    std::vector<pf_pos48> kids((size_t)nm);
    for (int i = 0; i < nm; i++) {
        kids[(size_t)i] = to_pos48(make_move(b, mv[i]), !black);
        moves[i] = (uint16_t)to_abs_move(mv[i], black);
    }
    *n = nm;
    if (!nm) return PF_OK;
    return pf_perft_batch(ctx, kids.data(), nm, depth - 1, counts);
}

pf_status pf_move_to_uci(uint16_t m, char* buf) {
    if (!buf) return PF_EINVAL;
    const int from = m & 63, to = (m >> 6) & 63;
    buf[0] = (char)('a' + (from & 7));
    buf[1] = (char)('1' + (from >> 3));
    buf[2] = (char)('a' + (to & 7));
    buf[3] = (char)('1' + (to >> 3));
    const int pc = promo_code(m);
    buf[4] = pc ? "nbrq"[pc - 1] : 0;
    buf[5] = 0;
    return PF_OK;
}

pf_status pf_apply_uci(pf_pos48* pos, const char* uci) {
    if (!pos || !uci || strlen(uci) < 4) return PF_EINVAL;
    const int black = pos_black_to_move(*pos);
    const Board b = to_board(*pos);
    u32 mv[256];
    const int nm = host_moves(b, mv);
    for (int i = 0; i < nm; i++) {
        char t[6];
        pf_move_to_uci((uint16_t)to_abs_move(mv[i], black), t);
        if (!strncmp(t, uci, 4) && t[4] == (uci[4] == ' ' ? 0 : uci[4])) {
            *pos = to_pos48(make_move(b, mv[i]), !black);
            return PF_OK;
        }
    }
    return PF_EINVAL;
}

pf_status pf_set_backend(pf_ctx* ctx, int backend) {
    if (!ctx || backend < 0 || backend > 3) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    ctx->eng.backend = backend;
    return PF_OK;
}

pf_status pf_tune_mode(pf_ctx* ctx, const pf_pos48* sample, int n, int depth, int what, const char* path) {
    if (!ctx || !sample || n <= 0) return PF_EINVAL;
    if (what == PF_TUNE_ALL) what = PF_TUNE_SINGLE_PASS | PF_TUNE_TABLE;
    if (what & ~(PF_TUNE_SINGLE_PASS | PF_TUNE_TABLE)) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    if (ctx->srv) return PF_EBUSY;
    pf_status s = ctx->eng.autotune(sample, n, depth, what);
    if (s == PF_OK && path) s = ctx->eng.save_tuning(path);
    return s;
}

pf_status pf_tune(pf_ctx* ctx, const pf_pos48* sample, int n, int depth, const char* path) {
    return pf_tune_mode(ctx, sample, n, depth, PF_TUNE_ALL, path);
}

pf_status pf_set_exact_ep(pf_ctx* ctx, int on) {
    if (!ctx) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    if (ctx->eng.serving) return PF_EBUSY;
    ctx->eng.exact_ep = on != 0;
    return PF_OK;
}

pf_status pf_set_tt_regions(pf_ctx* ctx, int log2_regions) {
    if (!ctx || log2_regions < 0 || log2_regions > 4) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    if (ctx->eng.serving) return PF_EBUSY;
    if (!ctx->eng.d_tt) return PF_EINVAL;
    if ((ctx->eng.tt_buckets >> log2_regions) < 1024) return PF_EINVAL;
    ctx->eng.tt_split = log2_regions;
    ctx->eng.drop_plan_graphs();  // captured passes hold the table's mask word
    return ctx->eng.tt_clear();
}

pf_status pf_set_level_graphs(pf_ctx* ctx, int on) {
    if (!ctx) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    ctx->eng.tt_graphs = on != 0;
    return PF_OK;
}

pf_status pf_set_dedup(pf_ctx* ctx, int on) {
    if (!ctx) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    ctx->eng.dedup = on != 0;
    return PF_OK;
}

pf_status pf_tt_save(pf_ctx* ctx, const char* path) {
    if (!ctx || !path) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    if (ctx->eng.serving) return PF_EBUSY;
    return ctx->eng.tt_save(path);
}

pf_status pf_tt_load(pf_ctx* ctx, const char* path) {
    if (!ctx || !path) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    if (ctx->eng.serving) return PF_EBUSY;
    return ctx->eng.tt_load(path);
}

pf_status pf_get_stats(pf_ctx* ctx, pf_stats* out) {
    if (!ctx || !out) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    *out = ctx->eng.st;
    if (ctx->srv) {  // the serving loop keeps its own counters
        u64 steps = 0, segs = 0, dups = 0, probes = 0, hits = 0;
        double busy = 0;
        ctx->srv->counters(&steps, &segs, &busy, &dups, &probes, &hits);
        out->levels_run += steps;
        out->dup_nodes += dups;
        out->tt_probes = probes;
        out->tt_hits = hits;
        out->gpu_ms += busy * 1e3;  // step graphs, launch to completion
    }
    return PF_OK;
}

pf_status pf_serve_start(pf_ctx* ctx, int max_jobs, uint64_t job_bytes) {
    if (!ctx || max_jobs < 1) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    if (ctx->srv) return PF_EBUSY;
    if (ctx->eng.arena_top != 0) return PF_EBUSY;
    Server* srv = new (std::nothrow) Server(ctx->eng, max_jobs, job_bytes);
    if (!srv) return PF_ENOMEM;
    const pf_status s = srv->start();
    if (s != PF_OK) {
        delete srv;
        return s;
    }
    {
        std::lock_guard<std::mutex> gs(ctx->srv_mu);
        ctx->srv = srv;
    }
    ctx->eng.serving = true;
    return PF_OK;
}

pf_status pf_submit(pf_ctx* ctx, const pf_pos48* pos, int depth, uint64_t tag) {
    if (!ctx || !pos) return PF_EINVAL;
    std::lock_guard<std::mutex> g(ctx->srv_mu);
    if (!ctx->srv) return PF_EINVAL;
    return ctx->srv->submit(*pos, depth, tag);
}

int pf_poll(pf_ctx* ctx, pf_result* out, int max) {
    if (!ctx || !out || max < 0) return PF_EINVAL;
    std::lock_guard<std::mutex> g(ctx->srv_mu);
    if (!ctx->srv) return PF_EINVAL;
    return ctx->srv->poll(out, max);
}

pf_status pf_serve_stop(pf_ctx* ctx) {
    if (!ctx) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    if (!ctx->srv) return PF_OK;
    // drain first (pf_submit sees PF_EBUSY, pf_poll still answers), then
    // unpublish the server under srv_mu before deleting it
    ctx->srv->stop();
    pf::Server* s = nullptr;
    {
        std::lock_guard<std::mutex> gs(ctx->srv_mu);
        std::swap(s, ctx->srv);
    }
    delete s;
    ctx->eng.serving = false;
    return PF_OK;
}

double pf_now(void) { return now_s(); }

pf_status pf_sync(pf_ctx* ctx) {
    if (!ctx) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    if (hipSetDevice(ctx->eng.device) != hipSuccess || hipStreamSynchronize(ctx->eng.stream) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
        return PF_EDEVICE;
    return PF_OK;
}

pf_status pf_probe(pf_ctx* ctx, int kind, uint64_t bytes, uint64_t iters, double* out) {
    if (!ctx || !out) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    return ctx->eng.probe(kind, bytes, iters, out);
}

pf_status pf_reset_stats(pf_ctx* ctx) {
    if (!ctx) return PF_EINVAL;
    std::lock_guard<std::recursive_mutex> g(ctx->eng.mu);
    ctx->eng.st = pf_stats{};
    // the table counters accumulate on the device
    if (ctx->eng.d_ttstats &&
        (hipSetDevice(ctx->eng.device) != hipSuccess ||
         hipMemsetAsync(ctx->eng.d_ttstats, 0, 8 * sizeof(unsigned long long), ctx->eng.stream) != hipSuccess ||
         hipStreamSynchronize(ctx->eng.stream) != hipSuccess))
        return PF_EDEVICE;
    return PF_OK;
}

int pf_count_moves_host(const pf_pos48* pos) {
    if (!pos) return -1;
    return (int)count_moves(to_board(*pos), kHostSl);
}

pf_status pf_playouts(uint64_t seed, int n, int min_plies, int max_plies, int min_moves, int no_check,
                      pf_pos48* out) {
    if (n < 0 || (n && !out) || min_plies < 0 || max_plies < min_plies) return PF_EINVAL;
    u64 s = seed ? seed : 0x9E3779B97F4A7C15ULL;
    auto rng = [&s]() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; };
    pf_pos48 start;
    pf_fen_to_pos("rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1", &start);
    int got = 0;
    long attempts = 0;
    while (got < n) {
        if (++attempts > 1000000L + 1000L * n) return PF_EINVAL;
        const int L = min_plies + (int)(rng() % (u64)(max_plies - min_plies + 1));
        Board b = to_board(start);
        int black = 0;
        bool ok = true;
        for (int ply = 0; ply < L; ply++) {
            u32 mv[256];
            const int nm = host_moves(b, mv);
            if (!nm) { ok = false; break; }
            // canonical order: ascending (from, to, promotion) in absolute squares
            int key[256];
            for (int i = 0; i < nm; i++) {
                const u32 a = to_abs_move(mv[i], black);
                key[i] = (int)((a & 63) << 9 | ((a >> 6) & 63) << 3) | promo_code(a);
            }
            for (int i = 1; i < nm; i++) {
                const u32 m = mv[i];
                const int k = key[i];
                int j = i - 1;
                while (j >= 0 && key[j] > k) { mv[j + 1] = mv[j]; key[j + 1] = key[j]; j--; }
                mv[j + 1] = m;
                key[j + 1] = k;
            }
            b = make_move(b, mv[rng() % (u64)nm]);
            black ^= 1;
        }
        if (!ok) continue;
        PosInfo pi;
        analyse(b, kHostSl, pi);
        const int nm = (int)count_moves(b, kHostSl);
        if (nm == 0 || nm < min_moves) continue;
        if (no_check && pi.checkers) continue;
        out[got++] = to_pos48(b, black);
    }
    return PF_OK;
}

}  // extern "C"
