// engine.cpp -- host runtime: device set-up, frontier arena and the
// breadth-first level driver.
//
// Level driver (replaces perft_bb_gpu_simple's <<<1,1>>> CDP loop,
// perft_bb.h:1027-1138, and the host DFS above launchDepth,
// perft_bb.h:1143-1220 / launcher.h:1181-1389):
//   level 0 = the roots; a root kernel stores each root's legal-move count in
//   its Board.  For L = 0 .. depth-3 the expand kernel writes level L+1 and
//   returns the exact size of level L+2 (sum of the children's counts), so
//   every allocation is exact -- no (73 + 6*bf)*n*1.2 estimate
//   (perft_bb.h:1352-1363) and no out-of-memory retry (launcher.h:1049-1098).
//   When level L+1 does not fit in the arena, level L is cut into chunks and
//   each chunk is driven to the leaves before the next (depth-first chunking).
//   Level depth-2 is expanded in leaf mode: children are counted, not stored.
#include "engine.h"

#include <string.h>

#include <algorithm>

#include "tables.h"

namespace pf {

#define PF_HIP(x)                                     \
    do {                                              \
        if ((x) != hipSuccess) return PF_EDEVICE;     \
    } while (0)

static constexpr u64 kNodeBytes = sizeof(Board) + sizeof(u32);
static constexpr u64 kAlign = 256;

pf_status Engine::init(const pf_config& cfg) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PF_EDEVICE;
    if (cfg.device < 0 || cfg.device >= ndev) return PF_EINVAL;
    if (cfg.backend < 0 || cfg.backend > 3) return PF_EINVAL;
    device = cfg.device;
    backend = cfg.backend;
    PF_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    PF_HIP(hipGetDeviceProperties(&prop, device));
    n_cu = prop.multiProcessorCount;
    PF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));

    SliderTables* h = new SliderTables;
    const int rc = build_slider_tables(*h);
    if (rc) { delete h; return PF_EDEVICE; }
    if (hipMalloc(&d_tab, sizeof(SliderTables)) != hipSuccess) { delete h; return PF_ENOMEM; }
    const hipError_t ce = hipMemcpy(d_tab, h, sizeof(SliderTables), hipMemcpyHostToDevice);
    delete h;
    if (ce != hipSuccess) return PF_EDEVICE;

    size_t free_b = 0, total_b = 0;
    PF_HIP(hipMemGetInfo(&free_b, &total_b));
    u64 want = cfg.arena_bytes;
    if (!want) want = std::min<u64>((u64)(free_b * 0.6), 16ULL << 30);
    want = (want + kAlign - 1) & ~(kAlign - 1);
    if (hipMalloc(&arena, want) != hipSuccess) return PF_ENOMEM;
    arena_bytes = want;
    arena_top = 0;

    PF_HIP(hipMalloc(&d_ctrl, 64 * sizeof(u32)));
    PF_HIP(hipMalloc(&d_total, 8 * sizeof(unsigned long long)));
    PF_HIP(hipHostMalloc(&h_pinned, 64 * sizeof(unsigned long long), hipHostMallocDefault));
    PF_HIP(hipEventCreate(&ev0));
    PF_HIP(hipEventCreate(&ev1));
    PF_HIP(hipMalloc(&d_pctl, sizeof(PersistCtl)));
    for (int b = 0; b < 4; b++)
        for (int m = 0; m < 2; m++) {
            grid[b][m] = expand_grid(b, m, 0, n_cu);
            grid_small[b][m] = expand_grid(b, m, 1, n_cu);
        }
    apply_tuning();
    return ensure_sums(1024);
}

void Engine::destroy() {
    if (device >= 0) hipSetDevice(device);
    if (stream) hipStreamSynchronize(stream);
    hipFree(d_tab);
    hipFree(arena);
    hipFree(d_ctrl);
    hipFree(d_total);
    hipFree(d_sums);
    hipFree(d_wsums);
    hipFree(d_pctl);
    hipFree(d_tt);
    hipFree(d_z);
    hipFree(d_ttstats);
    d_tt = nullptr; d_z = nullptr; d_ttstats = nullptr; tt_buckets = 0;
    drop_plan_graphs();
    hipFree(d_plan);
    hipFree(d_ptab);
    hipHostFree(h_plan);
    hipHostFree(h_pstat);
    d_plan = nullptr; d_ptab = nullptr; h_plan = nullptr; h_pstat = nullptr;
    hipHostFree(h_pinned);
    if (ev0) hipEventDestroy(ev0);
    if (ev1) hipEventDestroy(ev1);
    if (stream) hipStreamDestroy(stream);
    d_tab = nullptr; arena = nullptr; d_ctrl = nullptr; d_total = nullptr;
    d_sums = nullptr; d_wsums = nullptr; h_pinned = nullptr; stream = nullptr; d_pctl = nullptr;
}

pf_status Engine::ensure_sums(u64 n) {
    if (n <= sums_cap) return PF_OK;
    u64 cap = std::max<u64>(n, sums_cap * 2);
    hipFree(d_sums);
    hipFree(d_wsums);
    d_sums = d_wsums = nullptr;
    if (hipMalloc(&d_sums, cap * 8) != hipSuccess) return PF_ENOMEM;
    if (hipMalloc(&d_wsums, cap * 8) != hipSuccess) return PF_ENOMEM;
    sums_cap = cap;
    return PF_OK;
}

char* Engine::alloc(u64 bytes) {
    bytes = (bytes + kAlign - 1) & ~(kAlign - 1);
    if (arena_top + bytes > arena_bytes) return nullptr;
    char* p = arena + arena_top;
    arena_top += bytes;
    st.peak_arena_bytes = std::max<u64>(st.peak_arena_bytes, arena_top);
    return p;
}

void Engine::release(char* p) { arena_top = (u64)(p - arena); }

pf_status Engine::range_count(const Board* b, u64 n, u64* total) {
    PF_HIP(hipMemsetAsync(d_total, 0, 8, stream));
    PF_HIP(launch_range_count(b, (u32)n, d_total, stream));
    PF_HIP(hipMemcpyAsync(h_pinned, d_total, 8, hipMemcpyDeviceToHost, stream));
    PF_HIP(hipStreamSynchronize(stream));
    *total = h_pinned[0];
    return PF_OK;
}

pf_status Engine::expand(int mode, const Board* in, const u32* in_root, u64 n_in, Board* out, u32* out_root,
                         u64 out_cap, unsigned long long* sums, u64* next_total) {
    if (n_in == 0) { if (next_total) *next_total = 0; return PF_OK; }
    if (n_in > 0xFFFFFFFFull) return PF_ENOMEM;
    PF_HIP(hipMemsetAsync(d_ctrl, 0, 4 * sizeof(u32), stream));
    if (next_total) PF_HIP(hipMemsetAsync(d_total, 0, 8, stream));
    LevelArgs a;
    a.in = in;
    a.in_root = in_root;
    a.n_in = (u32)n_in;
    a.out = out;
    a.out_root = out_root;
    a.out_cap = out_cap;
    a.out_count = d_ctrl + 2;
    a.root_sums = sums;
    a.total = next_total ? d_total : nullptr;
    a.tile_ctr = d_ctrl + 1;
    a.err = d_ctrl + 0;
    // a level with fewer than 64 parents per workgroup of the wide launch is
    // latency-bound: it takes the small-level shape
    const int small = n_in < (u64)grid[backend][mode] * 64 ? 1 : 0;
    int g = small ? grid_small[backend][mode] : grid[backend][mode];
    a.tile_parents = level_tile_parents(n_in, (u64)g, 256);
    const u64 tiles = (n_in + a.tile_parents - 1) / a.tile_parents;
    if ((u64)g > tiles) g = (int)tiles;
    if (timing) PF_HIP(hipEventRecord(ev0, stream));
    PF_HIP(launch_expand(backend, mode, small, DevTables{d_tab}, a, g, stream));
    if (timing) PF_HIP(hipEventRecord(ev1, stream));
    st.kernel_launches++;
    st.levels_run++;
    if (mode == kWrite) {
        PF_HIP(hipMemcpyAsync(h_pinned, d_ctrl, 4 * sizeof(u32), hipMemcpyDeviceToHost, stream));
        if (next_total) PF_HIP(hipMemcpyAsync(h_pinned + 4, d_total, 8, hipMemcpyDeviceToHost, stream));
        PF_HIP(hipStreamSynchronize(stream));
        const u32* c = reinterpret_cast<const u32*>(h_pinned);
        if (c[0]) return PF_ENOMEM;
        if (next_total) *next_total = h_pinned[4];
    }
    if (timing) {
        PF_HIP(hipEventSynchronize(ev1));
        float ms = 0;
        PF_HIP(hipEventElapsedTime(&ms, ev0, ev1));
        st.gpu_ms += ms;
        run_expand_ms += ms;
        run_expand_launches++;
    }
    return PF_OK;
}

// Drive `lvl` (level L, n nodes whose children number n_next) to the leaves.
pf_status Engine::descend(const Board* lvl, const u32* rts, u64 n, u64 n_next, int L) {
    if (n == 0) return PF_OK;
    if (depth - L == 2) return expand(kLeaf, lvl, rts, n, nullptr, nullptr, 0, d_sums, nullptr);
    const u64 need = n_next * kNodeBytes + 2 * kAlign;
    const u64 free_b = arena_bytes - arena_top;
    if (need <= free_b && n_next <= 0xFFFFFFFFull) {
        char* mark = arena + arena_top;
        Board* nb = reinterpret_cast<Board*>(alloc(n_next * sizeof(Board)));
        u32* nr = reinterpret_cast<u32*>(alloc(n_next * sizeof(u32)));
        if (!nb || !nr) { release(mark); return PF_ENOMEM; }
        u64 n_next2 = 0;
        pf_status s = expand(kWrite, lvl, rts, n, nb, nr, n_next, nullptr, &n_next2);
        if (s != PF_OK) { release(mark); return s; }
        s = descend(nb, nr, n_next, n_next2, L + 1);
        release(mark);
        return s;
    }
    // Depth-first chunking: split this level so each piece's children fit.
    if (n == 1) return PF_ENOMEM;  // a single node's children (<= 218) always fit a sane arena
    st.chunks++;
    const u64 pieces = std::max<u64>(2, (need + free_b - 1) / std::max<u64>(free_b / 2, 1) + 1);
    const u64 step = (n + pieces - 1) / pieces;
    for (u64 a = 0; a < n; a += step) {
        const u64 m = std::min(step, n - a);
        u64 nn = 0;
        pf_status s = range_count(lvl + a, m, &nn);
        if (s != PF_OK) return s;
        s = descend(lvl + a, rts + a, m, nn, L);
        if (s != PF_OK) return s;
    }
    return PF_OK;
}

pf_status Engine::upload_roots(const pf_pos48* pos, int n, Board** roots, u32** rid, bool dev_in) {
    *roots = reinterpret_cast<Board*>(alloc((u64)n * sizeof(Board)));
    *rid = reinterpret_cast<u32*>(alloc((u64)n * sizeof(u32)));
    if (!*roots || !*rid) return PF_ENOMEM;
    if (dev_in) {  // already resident: convert in place on the stream, no host round trip
        if (launch_pos_to_board(pos, (u32)n, *roots, *rid, stream) != hipSuccess) return PF_EDEVICE;
        st.kernel_launches++;
        PF_HIP(hipMemsetAsync(d_sums, 0, (size_t)n * 8, stream));
        return PF_OK;
    }
    std::vector<Board> hb((size_t)n);
    std::vector<u32> hr((size_t)n);
    for (int i = 0; i < n; i++) { hb[(size_t)i] = to_board(pos[i]); hr[(size_t)i] = (u32)i; }
    PF_HIP(hipMemcpyAsync(*roots, hb.data(), (size_t)n * sizeof(Board), hipMemcpyHostToDevice, stream));
    PF_HIP(hipMemcpyAsync(*rid, hr.data(), (size_t)n * sizeof(u32), hipMemcpyHostToDevice, stream));
    PF_HIP(hipMemsetAsync(d_sums, 0, (size_t)n * 8, stream));
    PF_HIP(hipStreamSynchronize(stream));  // host vectors go out of scope
    return PF_OK;
}

pf_status Engine::perft_levels(const pf_pos48* pos, int n, int d, u64* out) {
    if (n <= 0) return PF_OK;
    if (d <= 0) {
        for (int i = 0; i < n; i++) out[i] = 1;
        return PF_OK;
    }
    if (d_tt && d >= 3) return perft_levels_tt(pos, n, d, out);
    PF_HIP(hipSetDevice(device));
    pf_status s = ensure_sums((u64)n);
    if (s != PF_OK) return s;
    depth = d;
    run_expand_ms = 0;
    run_expand_launches = 0;
    char* mark = arena + arena_top;
    auto fail = [&](pf_status e) { release(mark); return e; };
    Board* roots = nullptr;
    u32* rid = nullptr;
    s = upload_roots(pos, n, &roots, &rid);
    if (s != PF_OK) return fail(s);
    // root move counts; perft(1) lands in d_sums
    if (launch_init_roots(backend, DevTables{d_tab}, roots, (u32)n, d_sums, stream) != hipSuccess)
        return fail(PF_EDEVICE);
    st.kernel_launches++;
    if (d >= 2) {
        u64 n1 = 0;
        if (d >= 3) {
            s = range_count(roots, (u64)n, &n1);
            if (s != PF_OK) return fail(s);
        }
        if (hipMemsetAsync(d_sums, 0, (size_t)n * 8, stream) != hipSuccess) return fail(PF_EDEVICE);
        s = descend(roots, rid, (u64)n, n1, 0);
        if (s != PF_OK) return fail(s);
    }
    if (hipMemcpyAsync(out, d_sums, (size_t)n * 8, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return fail(PF_EDEVICE);
    st.last_expand_ms = run_expand_launches ? run_expand_ms / (double)run_expand_launches : 0.0;
    st.last_expand_launches = run_expand_launches;
    release(mark);
    return PF_OK;
}

pf_status Engine::perft_stream(const pf_pos48* pos, int n, int d, u64* out, bool dev_in, bool dev_out) {
    if (n <= 0) return PF_OK;
    if (d <= 1) {
        if (dev_in || dev_out) return PF_EINVAL;  // callers stage shallow device batches themselves
        return perft_levels(pos, n, d, out);
    }
    PF_HIP(hipSetDevice(device));
    pf_status s = ensure_sums((u64)n);
    if (s != PF_OK) return s;
    char* mark = arena + arena_top;
    auto fail = [&](pf_status e) { release(mark); return e; };
    Board* roots = nullptr;
    u32* rid = nullptr;
    s = upload_roots(pos, n, &roots, &rid, dev_in);
    if (s != PF_OK) return fail(s);
    if (hipMemsetAsync(d_pctl, 0, sizeof(PersistCtl), stream) != hipSuccess) return fail(PF_EDEVICE);
    PersistArgs p;
    p.roots = roots;
    p.root_ids = rid;
    p.n_roots = (u32)n;
    p.depth = d;
    p.arena = arena + arena_top;
    p.arena_bytes = arena_bytes - arena_top;
    p.ctl = d_pctl;
    p.sums = d_sums;
    if (timing && hipEventRecord(ev0, stream) != hipSuccess) return fail(PF_EDEVICE);
    const int var[2] = {tune[backend].s_leaf, tune[backend].s_write};
    if (launch_stream(backend, DevTables{d_tab}, p, sgrid[backend], var, stream) != hipSuccess) return fail(PF_EDEVICE);
    if (timing && hipEventRecord(ev1, stream) != hipSuccess) return fail(PF_EDEVICE);
    st.kernel_launches += (u64)d;  // the root count and d - 1 levels
    st.levels_run += (u64)(d - 1);
    PersistCtl h;
    if (hipMemcpyAsync(&h, d_pctl, sizeof(PersistCtl), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipMemcpyAsync(out, d_sums, (size_t)n * 8, dev_out ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                       stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return fail(PF_EDEVICE);
    if (timing) {
        float ms = 0;
        hipEventElapsedTime(&ms, ev0, ev1);
        st.gpu_ms += ms;
        st.last_expand_ms = ms;
        st.last_expand_launches = 1;
    }
    release(mark);
    if (h.err & 2) return PF_ENOMEM;
    if (h.err) return PF_EDEVICE;
    return PF_OK;
}

pf_status Engine::perft_persistent(const pf_pos48* pos, int n, int d, u64* out, bool dev_in, bool dev_out) {
    if (n <= 0) return PF_OK;
    if (d <= 1) {
        if (dev_in || dev_out) return PF_EINVAL;  // callers stage shallow device batches themselves
        return perft_levels(pos, n, d, out);
    }
    PF_HIP(hipSetDevice(device));
    pf_status s = ensure_sums((u64)n);
    if (s != PF_OK) return s;
    char* mark = arena + arena_top;
    auto fail = [&](pf_status e) { release(mark); return e; };
    Board* roots = nullptr;
    u32* rid = nullptr;
    s = upload_roots(pos, n, &roots, &rid, dev_in);
    if (s != PF_OK) return fail(s);
    PF_HIP(hipMemsetAsync(d_pctl, 0, sizeof(PersistCtl), stream));
    PersistArgs p;
    p.roots = roots;
    p.root_ids = rid;
    p.n_roots = (u32)n;
    p.depth = d;
    p.arena = arena + arena_top;
    p.arena_bytes = arena_bytes - arena_top;
    p.ctl = d_pctl;
    p.sums = d_sums;
    if (timing && hipEventRecord(ev0, stream) != hipSuccess) return fail(PF_EDEVICE);
    if (launch_persistent(backend, tune[backend].p_variant, DevTables{d_tab}, p, pgrid[backend], stream) !=
        hipSuccess) {
        // refused before it started (nothing ran): clear the error for the caller's fallback
        (void)hipGetLastError();
        return fail(PF_EUNAVAIL);
    }
    if (timing && hipEventRecord(ev1, stream) != hipSuccess) return fail(PF_EDEVICE);
    st.kernel_launches++;
    PersistCtl h;
    if (hipMemcpyAsync(&h, d_pctl, sizeof(PersistCtl), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipMemcpyAsync(out, d_sums, (size_t)n * 8, dev_out ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                       stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return fail(PF_EDEVICE);
    if (timing) {
        float ms = 0;
        hipEventElapsedTime(&ms, ev0, ev1);
        st.gpu_ms += ms;
        st.last_expand_ms = ms;
        st.last_expand_launches = 1;
    }
    release(mark);
    if (h.err & 2) return PF_ENOMEM;
    if (h.err) return PF_EDEVICE;
    return PF_OK;
}


}  // namespace pf
