// tune.cpp -- the launch-parameter autotuner and its file format.
//
// The reference fixes its launch shape at compile time (BLOCK_SIZE and
// __launch_bounds__, perft_bb.h:327-329, switches.h:14-21) and picks its
// launch depth with a floating-point estimate (launcher.h:1535-1563).  Here
// the persistent kernel's tile shape and workgroups per CU, and the TT expand
// kernel's workgroups per CU, are chosen per slider backend by timing a
// caller-supplied sample on the device (the results are the same for every
// choice; only the time differs), and kept in a small text file:
//
//   perft_gpu_amd tuning 1
//   cus 256
//   backend ks persistent_variant 0 persistent_blocks_per_cu 3 tt_blocks_per_cu 3
//   ...
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <vector>

#include "engine.h"

namespace pf {

static const char* const kBackendName[4] = {"ks", "plain", "fancy", "byte"};
static const char* const kSinglePassName[2] = {"stream", "persistent"};

void Engine::apply_tuning() {
    drop_plan_graphs();  // captured passes hold the old launch grids
    for (int b = 0; b < 4; b++) {
        Tune& t = tune[b];
        if (t.p_variant < 0 || t.p_variant >= kPersistVariants) t.p_variant = 0;
        const bool lds_form_ok = b == kByte;
        if (t.s_leaf < 0 || (t.s_leaf >= kStreamVariants && !(lds_form_ok && t.s_leaf == kStreamByteLds)))
            t.s_leaf = kStreamLeafDefault;
        if (t.s_write < 0 || (t.s_write >= kStreamWriteVariants && !(lds_form_ok && t.s_write == kStreamByteLds)))
            t.s_write = kStreamWriteDefault;
        if (t.tt_chunk_pct < 10 || t.tt_chunk_pct > 90) t.tt_chunk_pct = 50;
        const int pmax = persistent_max_bpc(b, t.p_variant);
        const int pb = (t.p_bpc > 0 && t.p_bpc < pmax) ? t.p_bpc : pmax;
        pgrid[b] = pb * n_cu;
        sgrid[b][kLeaf] = stream_grid(b, kLeaf, t.s_leaf, n_cu);
        sgrid[b][kWrite] = stream_grid(b, kWrite, t.s_write, n_cu);
        if (d_tt) {
            const int tmax = expand_tt_max_bpc(b);
            const int tb = (t.tt_bpc > 0 && t.tt_bpc < tmax) ? t.tt_bpc : tmax;
            tt_grid[b] = tb * n_cu;
            const int lmax = expand_tt_leaf_bpc(b);
            const int lb = (t.tt_leaf_bpc > 0 && t.tt_leaf_bpc < lmax) ? t.tt_leaf_bpc : lmax;
            tt_leaf_grid[b] = lb * n_cu;
        }
    }
}

// One "backend <name> <key> <value> ..." line; keys may come in any order and
// any may be missing (an older file lacks the later ones: defaults).
static pf_status parse_backend_line(const char* line, int* backend, Engine::Tune* t) {
    char buf[512];
    strncpy(buf, line, sizeof buf - 1);
    buf[sizeof buf - 1] = 0;
    char* save = nullptr;
    char* tok = strtok_r(buf, " \t\r\n", &save);
    if (!tok || strcmp(tok, "backend") != 0) return PF_EINVAL;
    tok = strtok_r(nullptr, " \t\r\n", &save);
    *backend = -1;
    for (int b = 0; tok && b < 4; b++)
        if (strcmp(tok, kBackendName[b]) == 0) *backend = b;
    *t = Engine::Tune{};
    for (;;) {
        char* key = strtok_r(nullptr, " \t\r\n", &save);
        if (!key) break;
        char* val = strtok_r(nullptr, " \t\r\n", &save);
        if (!val) return PF_EINVAL;
        if (strcmp(key, "single_pass") == 0) {
            if (strcmp(val, kSinglePassName[0]) == 0) t->single_pass = 0;
            else if (strcmp(val, kSinglePassName[1]) == 0) t->single_pass = 1;
            else return PF_EINVAL;
            continue;
        }
        char* end = nullptr;
        const long v = strtol(val, &end, 10);
        if (!end || *end) return PF_EINVAL;
        if (strcmp(key, "persistent_variant") == 0) t->p_variant = (int)v;
        else if (strcmp(key, "persistent_blocks_per_cu") == 0) t->p_bpc = (int)v;
        else if (strcmp(key, "tt_blocks_per_cu") == 0) t->tt_bpc = (int)v;
        else if (strcmp(key, "stream_leaf") == 0) t->s_leaf = (int)v;
        else if (strcmp(key, "stream_write") == 0) t->s_write = (int)v;
        else if (strcmp(key, "tt_leaf_blocks_per_cu") == 0) t->tt_leaf_bpc = (int)v;
        else if (strcmp(key, "tt_chunk_pct") == 0) t->tt_chunk_pct = (int)v;
        else return PF_EINVAL;
    }
    return PF_OK;
}

pf_status Engine::load_tuning(const char* path) {
    FILE* f = fopen(path, "r");
    if (!f) return PF_OK;  // no file yet: defaults
    char line[512];
    pf_status s = PF_OK;
    if (!fgets(line, sizeof line, f) || strncmp(line, "perft_gpu_amd tuning 1", 22) != 0) s = PF_EINVAL;
    while (s == PF_OK && fgets(line, sizeof line, f)) {
        if (strncmp(line, "backend ", 8) != 0) continue;
        int b = -1;
        Tune t;
        s = parse_backend_line(line, &b, &t);
        if (s == PF_OK && b >= 0) tune[b] = t;
    }
    fclose(f);
    apply_tuning();
    return s;
}

pf_status Engine::save_tuning(const char* path) const {
    FILE* f = fopen(path, "w");
    if (!f) return PF_EINVAL;
    fprintf(f, "perft_gpu_amd tuning 1\ncus %d\n", n_cu);
    for (int b = 0; b < 4; b++)
        fprintf(f,
                "backend %s persistent_variant %d persistent_blocks_per_cu %d tt_blocks_per_cu %d single_pass %s "
                "stream_leaf %d stream_write %d tt_leaf_blocks_per_cu %d tt_chunk_pct %d\n",
                kBackendName[b], tune[b].p_variant, tune[b].p_bpc, tune[b].tt_bpc,
                kSinglePassName[tune[b].single_pass ? 1 : 0], tune[b].s_leaf, tune[b].s_write, tune[b].tt_leaf_bpc,
                tune[b].tt_chunk_pct);
    return fclose(f) == 0 ? PF_OK : PF_EINVAL;
}

// Times the current backend on the sample (positions, depth).
//  kTuneSinglePass: every tile shape x workgroups-per-CU of the cooperative
//    kernel, then the stream-ordered form against the best of them, then the
//    stream-ordered leaf-level kernel's forms (tile shape x waves per SIMD)
//    and, with the best leaf form, its write-level kernel's forms; kernel
//    time from HIP events, best of three after a warm-up.
//  kTuneTable (with a table): the level driver's write-level and leaf-level
//    workgroups per CU and, when the sample makes the driver chunk, the share
//    of the free arena a chunk's next level may take; whole-call time, best
//    of two after a warm-up, table cleared before each run.
// Every choice must give the same counts.
pf_status Engine::autotune(const pf_pos48* pos, int n, int d, int what) {
    if (!pos || n <= 0 || d < 2 || d > 16) return PF_EINVAL;
    std::vector<u64> out((size_t)n), want((size_t)n);
    const int b = backend;
    const Tune keep = tune[b];
    pf_status s = PF_OK;
    bool have_want = false;
    auto same = [&](const std::vector<u64>& got) -> bool {
        if (!have_want) {
            want = got;
            have_want = true;
            return true;
        }
        return got == want;  // launch shapes must never change a count
    };
    auto time_pass = [&](bool cooperative, double* ms) -> pf_status {
        double best = 1e30;
        for (int r = 0; r < 4; r++) {
            pf_status e = cooperative ? perft_persistent(pos, n, d, out.data()) : perft_stream(pos, n, d, out.data());
            if (e != PF_OK) return e;
            if (!same(out)) return PF_EDEVICE;
            if (r > 0 && st.last_expand_ms < best) best = st.last_expand_ms;
        }
        *ms = best;
        return PF_OK;
    };
    Tune best = keep;
    if (what & kTuneSinglePass) {
        double best_ms = 1e30;
        for (int v = 0; v < kPersistVariants && s == PF_OK; v++) {
            tune[b].p_variant = v;
            tune[b].p_bpc = 0;
            apply_tuning();
            const int pmax = persistent_max_bpc(b, v);
            for (int pb = pmax; pb >= 1 && pb >= pmax - 2; pb--) {
                tune[b].p_bpc = pb;
                apply_tuning();
                double ms = 0;
                s = time_pass(true, &ms);
                if (s == PF_ENOMEM || s == PF_EUNAVAIL) {  // the sample does not fit, or no cooperative launch now
                    s = PF_OK;
                    break;
                }
                if (s != PF_OK) break;
                if (ms < best_ms) {
                    best_ms = ms;
                    best.p_variant = v;
                    best.p_bpc = pb;
                }
            }
        }
        tune[b].p_variant = best.p_variant;
        tune[b].p_bpc = best.p_bpc;
        apply_tuning();
        if (s == PF_OK) {  // the stream-ordered form: its leaf kernel's forms, then its write kernel's
            double stream_ms = 1e30;
            for (int pass = 0; pass < 2 && s == PF_OK; pass++) {
                // the forms of this mode, and for the byte backend its LDS form
                const int nsv = (pass == 0 ? kStreamVariants : kStreamWriteVariants) + (b == kByte ? 1 : 0);
                for (int i = 0; i < nsv; i++) {
                    const int v = i < (pass == 0 ? kStreamVariants : kStreamWriteVariants) ? i : kStreamByteLds;
                    if (pass == 0) tune[b].s_leaf = v; else tune[b].s_write = v;
                    apply_tuning();
                    double ms = 0;
                    s = time_pass(false, &ms);
                    if (s != PF_OK) break;
                    if (ms < stream_ms) {
                        stream_ms = ms;
                        best.s_leaf = tune[b].s_leaf;
                        best.s_write = tune[b].s_write;
                    }
                }
                tune[b].s_leaf = best.s_leaf;
                tune[b].s_write = best.s_write;
                apply_tuning();
            }
            if (s == PF_ENOMEM) {  // does not fit unchunked: neither does the cooperative launch
                s = PF_OK;
                best.s_leaf = keep.s_leaf;
                best.s_write = keep.s_write;
            } else if (s == PF_OK) {
                best.single_pass = stream_ms <= best_ms ? 0 : 1;
            }
        }
        tune[b].s_leaf = best.s_leaf;
        tune[b].s_write = best.s_write;
        tune[b].single_pass = best.single_pass;
        apply_tuning();
    }
    if (s == PF_OK && (what & kTuneTable) && d_tt && d >= 3) {
        std::vector<u64> tout((size_t)n);
        bool chunked = false;
        auto time_tt = [&](double* ms) -> pf_status {
            double t_best = 1e30;
            for (int r = 0; r < 3; r++) {
                pf_status e = tt_clear();
                if (e != PF_OK) return e;
                const u64 c0 = st.chunks;
                const auto t0 = std::chrono::steady_clock::now();
                e = perft_levels(pos, n, d, tout.data());
                const double el =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (e != PF_OK) return e;
                if (!same(tout)) return PF_EDEVICE;
                if (st.chunks > c0) chunked = true;
                if (r > 0 && el < t_best) t_best = el;
            }
            *ms = t_best;
            return PF_OK;
        };
        // write-level workgroups per CU, then leaf-level, then the chunk share
        for (int knob = 0; knob < 3 && s == PF_OK; knob++) {
            double best_t = 1e30;
            int choices[3] = {0, 0, 0}, nc = 0;
            if (knob == 0) {
                const int tmax = expand_tt_max_bpc(b);
                for (int k = 0; k < 3 && tmax - k >= 1; k++) choices[nc++] = tmax - k;
            } else if (knob == 1) {
                const int lmax = expand_tt_leaf_bpc(b);
                for (int k = 0; k < 3 && lmax - k >= 1; k++) choices[nc++] = lmax - k;
            } else {
                if (!chunked) break;  // the sample never chunks: nothing to choose
                choices[0] = 30, choices[1] = 50, choices[2] = 70, nc = 3;
            }
            for (int c = 0; c < nc && s == PF_OK; c++) {
                if (knob == 0) tune[b].tt_bpc = choices[c];
                else if (knob == 1) tune[b].tt_leaf_bpc = choices[c];
                else tune[b].tt_chunk_pct = choices[c];
                apply_tuning();
                double ms = 0;
                s = time_tt(&ms);
                if (s == PF_OK && ms < best_t) {
                    best_t = ms;
                    if (knob == 0) best.tt_bpc = choices[c];
                    else if (knob == 1) best.tt_leaf_bpc = choices[c];
                    else best.tt_chunk_pct = choices[c];
                }
            }
            tune[b].tt_bpc = best.tt_bpc;
            tune[b].tt_leaf_bpc = best.tt_leaf_bpc;
            tune[b].tt_chunk_pct = best.tt_chunk_pct;
            apply_tuning();
        }
        if (s == PF_OK) s = tt_clear();
    }
    if (s != PF_OK) tune[b] = keep;
    apply_tuning();
    return s;
}

}  // namespace pf
