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
    for (int b = 0; b < 4; b++) {
        Tune& t = tune[b];
        if (t.p_variant < 0 || t.p_variant >= kPersistVariants || b == kByte) t.p_variant = 0;
        const int pmax = persistent_max_bpc(b, t.p_variant);
        const int pb = (t.p_bpc > 0 && t.p_bpc < pmax) ? t.p_bpc : pmax;
        pgrid[b] = pb * n_cu;
        if (d_tt) {
            const int tmax = expand_tt_max_bpc(b);
            const int tb = (t.tt_bpc > 0 && t.tt_bpc < tmax) ? t.tt_bpc : tmax;
            tt_grid[b] = tb * n_cu;
            tt_leaf_grid[b] = expand_tt_leaf_bpc(b) * n_cu;
        }
    }
}

pf_status Engine::load_tuning(const char* path) {
    FILE* f = fopen(path, "r");
    if (!f) return PF_OK;  // no file yet: defaults
    char line[256];
    pf_status s = PF_OK;
    if (!fgets(line, sizeof line, f) || strncmp(line, "perft_gpu_amd tuning 1", 22) != 0) s = PF_EINVAL;
    while (s == PF_OK && fgets(line, sizeof line, f)) {
        char name[16], sp[16] = "stream";
        int v = 0, pb = 0, tb = 0;
        if (strncmp(line, "backend ", 8) != 0) continue;
        const int got = sscanf(line,
                               "backend %15s persistent_variant %d persistent_blocks_per_cu %d tt_blocks_per_cu %d "
                               "single_pass %15s",
                               name, &v, &pb, &tb, sp);
        const int single = strcmp(sp, kSinglePassName[1]) == 0 ? 1 : strcmp(sp, kSinglePassName[0]) == 0 ? 0 : -1;
        if (got < 4 || single < 0) {
            s = PF_EINVAL;
            break;
        }
        for (int b = 0; b < 4; b++)
            if (strcmp(name, kBackendName[b]) == 0) tune[b] = Tune{v, pb, tb, single};
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
        fprintf(f, "backend %s persistent_variant %d persistent_blocks_per_cu %d tt_blocks_per_cu %d single_pass %s\n",
                kBackendName[b], tune[b].p_variant, tune[b].p_bpc, tune[b].tt_bpc,
                kSinglePassName[tune[b].single_pass ? 1 : 0]);
    return fclose(f) == 0 ? PF_OK : PF_EINVAL;
}

// Times the current backend on the sample (positions, depth): every tile shape
// x workgroups-per-CU of the persistent kernel and the stream-ordered level
// launches (kernel time from HIP events, best of three after a warm-up), then,
// with a table, the TT expand kernel's workgroups per CU (whole level-driver
// time, table cleared before each run).
pf_status Engine::autotune(const pf_pos48* pos, int n, int d) {
    if (!pos || n <= 0 || d < 2 || d > 16) return PF_EINVAL;
    std::vector<u64> out((size_t)n), want((size_t)n);
    const int b = backend;
    const Tune keep = tune[b];
    pf_status s = PF_OK;
    bool have_want = false;
    auto time_pass = [&](bool cooperative, double* ms) -> pf_status {
        double best = 1e30;
        for (int r = 0; r < 4; r++) {
            pf_status e = cooperative ? perft_persistent(pos, n, d, out.data()) : perft_stream(pos, n, d, out.data());
            if (e != PF_OK) return e;
            if (!have_want) {
                want = out;
                have_want = true;
            } else if (out != want) {
                return PF_EDEVICE;  // launch shapes must never change a count
            }
            if (r > 0 && st.last_expand_ms < best) best = st.last_expand_ms;
        }
        *ms = best;
        return PF_OK;
    };
    double best_ms = 1e30;
    Tune best = keep;
    const int nv = (b == kByte) ? 1 : kPersistVariants;
    for (int v = 0; v < nv && s == PF_OK; v++) {
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
    if (s == PF_OK) {  // the stream-ordered form against the best cooperative shape
        double ms = 0;
        s = time_pass(false, &ms);
        if (s == PF_ENOMEM) {  // does not fit unchunked: neither does the cooperative launch
            s = PF_OK;
        } else if (s == PF_OK) {
            best.single_pass = ms <= best_ms ? 0 : 1;
        }
    }
    tune[b].single_pass = best.single_pass;
    if (s == PF_OK && d_tt) {
        const int tmax = expand_tt_max_bpc(b);
        double best_t = 1e30;
        std::vector<u64> tout((size_t)n);
        for (int tb = tmax; tb >= 1 && tb >= tmax - 2 && s == PF_OK; tb--) {
            tune[b].tt_bpc = tb;
            apply_tuning();
            double t_best = 1e30;
            for (int r = 0; r < 3 && s == PF_OK; r++) {
                s = tt_clear();
                if (s != PF_OK) break;
                const auto t0 = std::chrono::steady_clock::now();
                s = perft_levels(pos, n, d, tout.data());
                const double ms =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                if (s == PF_OK && have_want && tout != want) s = PF_EDEVICE;
                if (s == PF_OK && r > 0 && ms < t_best) t_best = ms;
            }
            if (s == PF_OK && t_best < best_t) {
                best_t = t_best;
                best.tt_bpc = tb;
            }
        }
        tune[b].tt_bpc = best.tt_bpc;
        if (s == PF_OK) s = tt_clear();
    }
    if (s != PF_OK) tune[b] = keep;
    apply_tuning();
    return s;
}

}  // namespace pf
