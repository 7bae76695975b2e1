// engine_tt.cpp -- level driver with the transposition table.
//
// Downsweep: every expand-and-write level probes its children's buckets and
// stores only the misses, each with both keys and its parent's index -- the
// first child of each position only: later children of the same position in
// the level are listed as duplicates; the last interior level is expanded in
// leaf mode into per-node accumulators.  Upsweep (on the way back up, level
// by level and chunk by chunk): every finished node adds its exact count to
// its parent and is stored in the table at its remaining depth, and every
// duplicate adds its original's count to its own parent.  Hits never change
// a count: the stored value is the node's exact perft, verified by a second
// 64-bit key.  This replaces the reference's per-depth table set
// (launcher.h:62-126, 155-320) and the hash kernels of perft_bb.h:664-961.
#include <string.h>

#include <algorithm>
#include <vector>

#include "engine.h"

namespace pf {

#define PF_HIP(x)                                 \
    do {                                          \
        if ((x) != hipSuccess) return PF_EDEVICE; \
    } while (0)

static constexpr u64 kTNodeBytes = sizeof(Board) + 4 + 8 + 8 + 4 + 8;

pf_status Engine::init_tt(u64 bytes, u64 seed) {
    zobrist_tables(h_z, seed);
    if (!bytes) return PF_OK;
    u64 buckets = 1;
    while (buckets * 2 * 32 <= bytes) buckets *= 2;
    if (buckets < 1024) return PF_EINVAL;
    PF_HIP(hipSetDevice(device));
    if (hipMalloc(&d_tt, buckets * 32) != hipSuccess) { d_tt = nullptr; return PF_ENOMEM; }
    tt_buckets = buckets;
    PF_HIP(hipMalloc(&d_z, sizeof(ZTables)));
    PF_HIP(hipMemcpy(d_z, &h_z, sizeof(ZTables), hipMemcpyHostToDevice));
    PF_HIP(hipMalloc(&d_ttstats, 8 * sizeof(unsigned long long)));
    PF_HIP(hipMemset(d_ttstats, 0, 8 * sizeof(unsigned long long)));
    // the planned passes' control block, step table and pinned parameters
    PF_HIP(hipMalloc(&d_plan, sizeof(TTPlanCtl)));
    PF_HIP(hipMalloc(&d_ptab, sizeof(StepTable)));
    PF_HIP(hipMemset(d_ptab, 0, sizeof(StepTable)));
    PF_HIP(hipHostMalloc(&h_plan, sizeof(TTPlanParams), hipHostMallocDefault));
    PF_HIP(hipHostMalloc(&h_pstat, sizeof(TTPlanStatus), hipHostMallocDefault));
    memset(h_plan, 0, sizeof(TTPlanParams));
    memset(h_pstat, 0, sizeof(TTPlanStatus));
    apply_tuning();
    return tt_clear();
}

void Engine::drop_plan_graphs() {
    for (int b = 0; b < 4; b++)
        for (int r = 0; r < kPlanMaxLevels; r++) {
            if (plan_exec[b][r]) hipGraphExecDestroy(plan_exec[b][r]);
            if (plan_graph[b][r]) hipGraphDestroy(plan_graph[b][r]);
            plan_exec[b][r] = nullptr;
            plan_graph[b][r] = nullptr;
        }
}

// The captured pass for chunks of remaining depth r with the current backend
// and launch grids (captured on first use, kept until the grids change).
pf_status Engine::plan_exec_for(int r, hipGraphExec_t* out) {
    if (r < 3 || r >= kPlanMaxLevels) return PF_EINVAL;
    if (!plan_exec[backend][r]) {
        const int wg = std::min(tt_grid[backend], step_max_bpc(backend, 0) * n_cu);
        const int lg = std::min(tt_leaf_grid[backend], step_max_bpc(backend, 1) * n_cu);
        if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return PF_EDEVICE;
        const hipError_t e = capture_plan_pass(backend, d_tab, d_z, d_plan, d_ptab, h_plan, h_pstat, d_tt,
                                               tt_mask_word(), d_ttstats, r, wg, lg, n_cu * 4, stream);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(stream, &g);
        if (e != hipSuccess || ee != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return PF_EDEVICE;
        }
        hipGraphExec_t x = nullptr;
        if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
            hipGraphDestroy(g);
            return PF_EDEVICE;
        }
        plan_graph[backend][r] = g;
        plan_exec[backend][r] = x;
    }
    *out = plan_exec[backend][r];
    return PF_OK;
}

// Predicted memory of a pass per child of its chunk: what the last passes of
// this remaining depth used, or before any, 112 B per written node (80 B of
// node, 32 B of duplicate table and list) over levels growing by the
// branching factor, plus a third.
double Engine::plan_bytes_per_child(int r, double bf) const {
    if (plan_bpc[r] > 0) return plan_bpc[r];
    const double b = std::min(48.0, std::max(8.0, bf));
    double sum = 0, lvl = 1;
    for (int k = 0; k <= r - 3; k++) {
        sum += lvl;
        lvl *= b;
    }
    return sum * 112.0 * 4.0 / 3.0 + 64.0;
}

// Parents [off, off + m) of level `lv` (level L; n_next children) as one
// replay of the captured pass.  PF_ENOMEM: it did not fit, and had no effect.
pf_status Engine::run_plan(const TLevel& lv, u64 off, u64 m, u64 n_next, int L) {
    const int r = depth - L;
    hipGraphExec_t ex = nullptr;
    pf_status s = plan_exec_for(r, &ex);
    if (s != PF_OK) return s;
    TTLevelRec top = {};
    top.b = lv.b + off;
    top.root = lv.root + off;
    top.k1 = lv.k1 + off;
    top.k2 = lv.k2 + off;
    top.acc = lv.acc + off;
    top.n = m;
    h_plan->top = top;
    h_plan->arena = arena + arena_top;
    h_plan->arena_bytes = arena_bytes - arena_top;
    h_plan->depth = r;
    h_plan->ply = L;
    h_plan->dedup = dedup ? 1 : 0;
    h_plan->ep_always = exact_ep ? 0 : 1;
    // the fused launch runs on the leaf grid the pass was captured with
    const int lg = std::min(tt_leaf_grid[backend], step_max_bpc(backend, 1) * n_cu);
    h_plan->fuse_min_tiles = (u32)(kFuseTilesPerWG * lg);
    if (timing) PF_HIP(hipEventRecord(ev0, stream));
    PF_HIP(hipGraphLaunch(ex, stream));
    if (timing) PF_HIP(hipEventRecord(ev1, stream));
    PF_HIP(hipStreamSynchronize(stream));
    const TTPlanStatus ps = *h_pstat;
    st.peak_arena_bytes = std::max<u64>(st.peak_arena_bytes, arena_top + ps.peak);
    const double used = (double)ps.peak / (double)std::max<u64>(n_next, 1);
    if (ps.err) {
        // the pass stopped before its writes reached anything outside its own
        // memory: predict more next time and let the caller redo this chunk
        plan_bpc[r] = std::max(plan_bpc[r], plan_bytes_per_child(r, (double)n_next / (double)m)) * 2.0;
        return PF_ENOMEM;
    }
    plan_bpc[r] = std::max(used * 1.15, plan_bpc[r] * 0.7);
    st.graph_passes++;
    st.kernel_launches += (u64)(5 * (r - 2) + 3);
    st.levels_run += (u64)(r - 1);
    st.made_nodes += ps.made;
    st.frontier_nodes += ps.frontier;
    st.dup_nodes += ps.dups;
    if (timing) {
        float ms = 0;
        PF_HIP(hipEventElapsedTime(&ms, ev0, ev1));
        st.gpu_ms += ms;
        run_expand_ms += ms;
        run_expand_launches++;
    }
    return PF_OK;
}

pf_status Engine::tt_clear() {
    if (!d_tt) return PF_OK;
    PF_HIP(hipMemsetAsync(d_tt, 0, tt_buckets * 32, stream));
    PF_HIP(hipStreamSynchronize(stream));
    return PF_OK;
}

bool Engine::alloc_tlevel(u64 n, TLevel* lv) {
    lv->b = reinterpret_cast<Board*>(alloc(n * sizeof(Board)));
    lv->root = reinterpret_cast<u32*>(alloc(n * 4));
    lv->k1 = reinterpret_cast<u64*>(alloc(n * 8));
    lv->k2 = reinterpret_cast<u64*>(alloc(n * 8));
    lv->parent = reinterpret_cast<u32*>(alloc(n * 4));
    lv->acc = reinterpret_cast<unsigned long long*>(alloc(n * 8));
    lv->bsum = nullptr;
    lv->dup = nullptr;
    lv->dup_n = 0;
    lv->dpar = lv->dslot = nullptr;
    lv->ndup = 0;
    return lv->b && lv->root && lv->k1 && lv->k2 && lv->parent && lv->acc;
}

pf_status Engine::expand_tt(int mode, const TLevel& lv, u64 off, u64 m, TLevel* nx, u64 cap, int L, u64* written,
                            u64* next_total) {
    if (!m) {
        if (written) *written = 0;
        if (next_total) *next_total = 0;
        return PF_OK;
    }
    if (m > 0xFFFFFFFFull || off + m > 0xFFFFFFFFull) return PF_ENOMEM;
    PF_HIP(hipMemsetAsync(d_ctrl, 0, 4 * sizeof(u32), stream));
    PF_HIP(hipMemsetAsync(d_total, 0, 8, stream));
    TTArgs a = {};
    a.in = lv.b + off;
    a.in_root = lv.root + off;
    a.in_k1 = lv.k1 + off;
    a.in_k2 = lv.k2 + off;
    a.n_in = (u32)m;
    a.in_base = (u32)off;
    a.in_acc = lv.acc + off;
    if (mode == kWrite) {
        a.out = nx->b;
        a.out_root = nx->root;
        a.out_k1 = nx->k1;
        a.out_k2 = nx->k2;
        a.out_parent = nx->parent;
        a.out_acc = nx->acc;
        a.out_cap = cap;
        a.total = d_total;
    }
    a.out_count = d_ctrl + 2;
    a.tt = d_tt;
    a.tt_mask = tt_mask_word();
    a.ep_always = exact_ep ? 0 : 1;
    a.child_depth = depth - (L + 1);
    a.mode = mode;
    a.tile_ctr = d_ctrl + 1;
    a.err = d_ctrl + 0;
    a.stats = d_ttstats;
    if (mode == kWrite && nx->dup) {
        a.dup = nx->dup;
        a.dup_n = nx->dup_n;
        a.dup_count = d_ctrl + 3;
        a.dup_par = nx->dpar;
        a.dup_slot = nx->dslot;
    }
    if (timing) PF_HIP(hipEventRecord(ev0, stream));
    PF_HIP(launch_expand_tt(backend, d_tab, d_z, a, mode == kLeaf ? tt_leaf_grid[backend] : tt_grid[backend], stream));
    if (timing) PF_HIP(hipEventRecord(ev1, stream));
    st.kernel_launches++;
    st.levels_run++;
    PF_HIP(hipMemcpyAsync(h_pinned, d_ctrl, 4 * sizeof(u32), hipMemcpyDeviceToHost, stream));
    PF_HIP(hipMemcpyAsync(h_pinned + 4, d_total, 8, hipMemcpyDeviceToHost, stream));
    PF_HIP(hipStreamSynchronize(stream));
    const u32* c = reinterpret_cast<const u32*>(h_pinned);
    if (c[0]) return PF_ENOMEM;
    if (written) *written = c[2];
    if (mode == kWrite) {
        st.frontier_nodes += c[2];
        st.dup_nodes += c[3];
        nx->ndup = c[3];
    }
    if (next_total) *next_total = h_pinned[4];
    if (timing) {
        float ms = 0;
        PF_HIP(hipEventElapsedTime(&ms, ev0, ev1));
        st.gpu_ms += ms;
        run_expand_ms += ms;
        run_expand_launches++;
    }
    return PF_OK;
}

// Slots of the duplicate table for a level of n nodes at ply `ply` (0: none).
static u64 dup_slots(bool on, u64 n, int ply) {
    if (!on || ply < kDupMinPly || n < 2) return 0;
    return dup_table_slots(n);
}

// Parents [off, off + m) of level `lv` (level L) have n_next children in all.
// Pieces whose predicted subtree fits the free arena run as planned passes
// (one graph replay each, no host round trip per level); what does not fit,
// or overflows its pass, takes the host-driven levels below.
pf_status Engine::descend_tt(const TLevel& lv, u64 off, u64 m, u64 n_next, int L) {
    if (!m) return PF_OK;
    const int r = depth - L;
    if (!tt_graphs || !d_plan || r < 3 || r >= kPlanMaxLevels) return descend_tt_host(lv, off, m, n_next, L);
    u64 a = off, left = m, left_next = n_next;
    bool split = false;
    while (left) {
        const u64 free_b = arena_bytes - arena_top;
        const double bpc = plan_bytes_per_child(r, (double)left_next / (double)left);
        const double fit = (double)free_b * 0.9 / bpc;  // children whose passes fit
        u64 len = left, nn = left_next;
        if ((double)left_next > fit) {
            len = std::max<u64>(1, (u64)((double)left * fit / (double)left_next));
            if (len < left) {
                pf_status s = range_count(lv.b + a, len, &nn);
                if (s != PF_OK) return s;
                if (!split) st.chunks++;  // this level runs depth-first in pieces
                split = true;
            }
        }
        pf_status s = PF_ENOMEM;
        if ((double)nn * bpc <= (double)free_b) s = run_plan(lv, a, len, nn, L);
        if (s == PF_ENOMEM) s = descend_tt_host(lv, a, len, nn, L);
        if (s != PF_OK) return s;
        a += len;
        left -= len;
        left_next -= nn;
    }
    return PF_OK;
}

// The host-driven form: write the next level (or chunk this one by the next
// level's size), descend, upsweep.
pf_status Engine::descend_tt_host(const TLevel& lv, u64 off, u64 m, u64 n_next, int L) {
    if (!m) return PF_OK;
    if (depth - L == 2) {
        st.made_nodes += n_next;
        return expand_tt(kLeaf, lv, off, m, nullptr, 0, L, nullptr, nullptr);
    }
    // the children's level, and for ply >= 3 their duplicate table and list,
    // live until the level's upsweep
    const u64 dn = dup_slots(dedup, n_next, L + 1);
    const u64 need = n_next * kTNodeBytes + (dn ? dn * sizeof(DupEntry) + n_next * 8 : 0) + 10 * kAlignTT;
    const u64 free_b = arena_bytes - arena_top;
    if (need <= free_b && n_next <= 0xFFFFFFFFull) {
        char* mark = arena + arena_top;
        TLevel nx;
        if (!alloc_tlevel(n_next, &nx)) { release(mark); return PF_ENOMEM; }
        if (dn) {
            nx.dup = reinterpret_cast<DupEntry*>(alloc(dn * sizeof(DupEntry)));
            nx.dup_n = dn;
            nx.dpar = reinterpret_cast<u32*>(alloc(n_next * 4));
            nx.dslot = reinterpret_cast<u32*>(alloc(n_next * 4));
            if (!nx.dup || !nx.dpar || !nx.dslot) { release(mark); return PF_ENOMEM; }
        }
        pf_status s = PF_OK;
        if (nx.dup && hipMemsetAsync(nx.dup, 0, dn * sizeof(DupEntry), stream) != hipSuccess) s = PF_EDEVICE;
        u64 written = 0, n_next2 = 0;
        st.made_nodes += n_next;
        if (s == PF_OK) s = expand_tt(kWrite, lv, off, m, &nx, n_next, L, &written, &n_next2);
        if (s == PF_OK) s = descend_tt(nx, 0, written, n_next2, L + 1);
        if (s == PF_OK && written + nx.ndup) {
            UpSeg u = {};
            u.parent = nx.parent;
            u.k1 = nx.k1;
            u.k2 = nx.k2;
            u.acc = nx.acc;
            u.parent_acc = lv.acc;
            u.dup = nx.dup;
            u.dup_par = nx.dpar;
            u.dup_slot = nx.dslot;
            u.n = (u32)written;
            u.ndup = (u32)nx.ndup;
            u.depth = depth - (L + 1);
            if (launch_upsweep(u, d_tt, tt_mask_word(), d_ttstats + 2, stream) != hipSuccess) s = PF_EDEVICE;
            st.kernel_launches++;
        }
        release(mark);
        return s;
    }
    if (m == 1) return PF_ENOMEM;
    st.chunks++;
    // each piece's next level may take the tuned share of the free arena
    const u64 budget = std::max<u64>(free_b / 100 * (u64)tune[backend].tt_chunk_pct, 1);
    const u64 pieces = std::max<u64>(2, (need + budget - 1) / budget + 1);
    const u64 step = (m + pieces - 1) / pieces;
    for (u64 a = off; a < off + m; a += step) {
        const u64 len = std::min(step, off + m - a);
        u64 nn = 0;
        pf_status s = range_count(lv.b + a, len, &nn);
        if (s != PF_OK) return s;
        s = descend_tt(lv, a, len, nn, L);
        if (s != PF_OK) return s;
    }
    return PF_OK;
}

pf_status Engine::perft_levels_tt(const pf_pos48* pos, int n, int d, u64* out) {
    PF_HIP(hipSetDevice(device));
    pf_status s = ensure_sums((u64)n);
    if (s != PF_OK) return s;
    depth = d;
    run_expand_ms = 0;
    run_expand_launches = 0;
    char* mark = arena + arena_top;
    auto fail = [&](pf_status e) { release(mark); return e; };
    TLevel r;
    if (!alloc_tlevel((u64)n, &r)) return fail(PF_ENOMEM);
    std::vector<Board> hb((size_t)n);
    std::vector<u32> hr((size_t)n);
    std::vector<u64> h1((size_t)n), h2((size_t)n);
    for (int i = 0; i < n; i++) {
        const int black = pos_black_to_move(pos[i]);
        hb[(size_t)i] = to_board(pos[i]);
        hr[(size_t)i] = (u32)i | ((u32)black << 31);
        h1[(size_t)i] = zobrist_full(h_z.k[0], hb[(size_t)i], black);
        h2[(size_t)i] = zobrist_full(h_z.k[1], hb[(size_t)i], black);
    }
    if (hipMemcpyAsync(r.b, hb.data(), (size_t)n * sizeof(Board), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipMemcpyAsync(r.root, hr.data(), (size_t)n * 4, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipMemcpyAsync(r.k1, h1.data(), (size_t)n * 8, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipMemcpyAsync(r.k2, h2.data(), (size_t)n * 8, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipMemsetAsync(d_total, 0, 8, stream) != hipSuccess)
        return fail(PF_EDEVICE);
    if (launch_roots_tt(backend, d_tab, r.b, r.k1, r.k2, (u32)n, d, d_tt, tt_mask_word(), r.acc, d_total, d_ttstats,
                        stream) != hipSuccess)
        return fail(PF_EDEVICE);
    st.kernel_launches++;
    if (hipMemcpyAsync(h_pinned + 4, d_total, 8, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return fail(PF_EDEVICE);
    const u64 n1 = h_pinned[4];
    s = descend_tt(r, 0, (u64)n, n1, 0);
    if (s != PF_OK) return fail(s);
    // the roots themselves, at full depth (answers repeated requests at once)
    UpSeg ru = {};
    ru.k1 = r.k1;
    ru.k2 = r.k2;
    ru.acc = r.acc;
    ru.n = (u32)n;
    ru.depth = d;
    if (launch_upsweep(ru, d_tt, tt_mask_word(), d_ttstats + 2, stream) != hipSuccess) return fail(PF_EDEVICE);
    unsigned long long ts[4];
    if (hipMemcpyAsync(out, r.acc, (size_t)n * 8, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipMemcpyAsync(ts, d_ttstats, sizeof(ts), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return fail(PF_EDEVICE);
    st.tt_probes = ts[0];
    st.tt_hits = ts[1];
    st.tt_stores = ts[2];
    st.last_expand_ms = run_expand_launches ? run_expand_ms / (double)run_expand_launches : 0.0;
    st.last_expand_launches = run_expand_launches;
    release(mark);
    return PF_OK;
}

}  // namespace pf
