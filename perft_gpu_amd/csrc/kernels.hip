// kernels.hip -- gfx950 kernels of the breadth-first perft.
//
// The reference runs one thread per frontier node in four kernels per level
// (make+count, moderngpu ScanD, IntervalExpand, generate; perft_bb.h:1071-1135)
// plus a leaf kernel (perft_bb.h:412-446).  Here one routine does a whole level:
//
//   tile of TP parents (from a global tile queue, so the grid is persistent)
//     -> block scan of the parents' move counts (stored in each Board)
//     -> each parent thread generates its legal moves into an LDS list
//     -> every thread of the workgroup takes (parent, move) entries from the
//        list: make the move, count the child's legal moves, and either
//          kLeaf : add the count to the parent's LDS sum (depth-1 nodes), or
//          kWrite: also append the child (count in its Board) to the next level
//     -> per-root sums: one 64-bit atomic per tile when the tile has a single
//        root (the common case), else one per parent; plus the level total.
//
// The LDS move list replaces the reference's global move array, its scan and
// its interval-expand (the parent index rides in the list entry), and the
// make+count is balanced across all 64 lanes regardless of how many moves each
// parent has.  Windows of CAP entries bound LDS use; a parent whose moves
// straddle a window boundary is generated once per window it touches.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include "level_dev.h"

namespace pf {


// The host-driven level loop has both shapes: the wide one above for large
// levels, and for small ones (few parents per workgroup, latency-bound) the
// 256 x 8192 tile without a waves-per-SIMD target, which does not spill.
template <int B, int MODE, int V>
__global__ __launch_bounds__(kWG)
__attribute__((amdgpu_waves_per_eu(V == LevelV<B>::v ? LevelWaves<B, MODE>::v : 1)))
void k_expand(const SliderTables* tabs, LevelArgs a) {
    __shared__ TileLds<PCfg<B, V>::TP, PCfg<B, V>::CAP, kWG> s;
    const auto sl = SlOf<B>::make(tabs);
    expand_level(sl, a, MODE, s);
}

template <int B>
__global__ __launch_bounds__(kWG) void k_init_roots(const SliderTables* t, Board* roots, u32 n,
                                                    unsigned long long* sums) {
    const typename SlOf<B>::T sl = SlOf<B>::make(t);
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const Board b = roots[i];
        const u32 c = count_moves(b, sl);
        roots[i] = with_count(b, c);
        sums[i] = c;
    }
}

// Device-resident roots in the reference layout -> side-to-move-frame boards
// and root ids (the batch may start from positions already in HBM, as the
// reference's perft_bb_gpu_simple_hash does).
__global__ __launch_bounds__(kWG) void k_pos_to_board(const pf_pos48* in, u32 n, Board* out, u32* rid) {
    const u32 i = blockIdx.x * kWG + threadIdx.x;
    if (i < n) {
        out[i] = to_board(in[i]);
        rid[i] = i;
    }
}

hipError_t launch_pos_to_board(const pf_pos48* in, u32 n, Board* out, u32* rid, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_pos_to_board, dim3((n + kWG - 1) / kWG), dim3(kWG), 0, s, in, n, out, rid);
    return hipGetLastError();
}

// Sum of the move counts stored in boards [0, n): the exact size of the next
// level for a chunk of this one.
__global__ __launch_bounds__(kWG) void k_range_count(const Board* b, u32 n, unsigned long long* out) {
    __shared__ unsigned long long s_red[kWG / 64];
    unsigned long long v = 0;
    for (u32 i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) v += (u64)meta_count(b[i]);
    v = wave_sum64(v);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kWG / 64; w++) t += s_red[w];
        if (t) atomicAdd(out, t);
    }
}

// One cooperative launch for a whole batch (replaces perft_bb_gpu_simple_hash's
// <<<1,1>>> level driver, perft_bb.h:1242-1589, and its ALLSET out-of-memory
// signal).  Every workgroup runs every level; grid barriers separate levels.
// All branch decisions between barriers read the same device words, so every
// workgroup takes the same path and reaches every barrier.
template <int B, int V>
__global__ __launch_bounds__(kWG) void k_persistent(const SliderTables* tabs, PersistArgs p) {
    namespace cg = cooperative_groups;
    cg::grid_group grid = cg::this_grid();
    __shared__ TileLds<PCfg<B, V>::TP, PCfg<B, V>::CAP, kWG> s;
    const auto sl = SlOf<B>::make(tabs);
    PersistCtl* ctl = p.ctl;

    // level 0: root move counts
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < p.n_roots; i += gridDim.x * blockDim.x) {
        const Board b = p.roots[i];
        const u32 c = count_moves(b, sl);
        p.roots[i] = with_count(b, c);
        if (p.depth == 1) p.sums[p.root_ids[i]] = c;
        atomicAdd(&ctl->totals[1], (unsigned long long)c);
    }
    if (p.depth <= 1) return;
    grid.sync();

    const Board* cur = p.roots;
    const u32* cur_r = p.root_ids;
    u64 n = p.n_roots, off = 0;
    for (int L = 0; L <= p.depth - 2; L++) {
        LevelArgs a;
        a.in = cur;
        a.in_root = cur_r;
        a.n_in = (u32)n;
        a.tile_ctr = &ctl->tile_ctr[L];
        a.err = &ctl->err;
        a.out_count = &ctl->out_count[L];
        a.tile_parents = level_tile_parents(n, gridDim.x, PCfg<B, V>::TP);
        if (L == p.depth - 2) {
            a.out = nullptr;
            a.out_root = nullptr;
            a.out_cap = 0;
            a.root_sums = p.sums;
            a.total = nullptr;
            expand_level(sl, a, kLeaf, s);
            break;
        }
        const u64 nn = __hip_atomic_load(&ctl->totals[L + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const u64 bbytes = (nn * sizeof(Board) + 255) & ~255ULL, rbytes = (nn * sizeof(u32) + 255) & ~255ULL;
        if (nn > 0xFFFFFFFFull || off + bbytes + rbytes > p.arena_bytes) {
            if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&ctl->err, 2u);
            break;  // uniform: every workgroup reads the same totals
        }
        Board* nb = reinterpret_cast<Board*>(p.arena + off);
        u32* nr = reinterpret_cast<u32*>(p.arena + off + bbytes);
        a.out = nb;
        a.out_root = nr;
        a.out_cap = nn;
        a.root_sums = nullptr;
        a.total = &ctl->totals[L + 2];
        expand_level(sl, a, kWrite, s);
        grid.sync();
        cur = nb;
        cur_r = nr;
        n = nn;
        off += bbytes + rbytes;
    }
}


// ---- launch plumbing --------------------------------------------------------
template <int B, int MODE>
static void* expand_fn(int small) {
    if (small) return reinterpret_cast<void*>(&k_expand<B, MODE, 0>);
    return reinterpret_cast<void*>(&k_expand<B, MODE, LevelV<B>::v>);
}

static void* pick_expand(int backend, int mode, int small) {
    switch (backend) {
    case kKS: return mode ? expand_fn<kKS, kWrite>(small) : expand_fn<kKS, kLeaf>(small);
    case kPlain: return mode ? expand_fn<kPlain, kWrite>(small) : expand_fn<kPlain, kLeaf>(small);
    case kFancy: return mode ? expand_fn<kFancy, kWrite>(small) : expand_fn<kFancy, kLeaf>(small);
    default: return mode ? expand_fn<kByte, kWrite>(small) : expand_fn<kByte, kLeaf>(small);
    }
}

template <int B>
static void* persistent_fn(int variant) {
    switch (variant) {
    case 1: return reinterpret_cast<void*>(&k_persistent<B, 1>);
    case 2: return reinterpret_cast<void*>(&k_persistent<B, 2>);
    default: return reinterpret_cast<void*>(&k_persistent<B, 0>);
    }
}

static void* pick_persistent(int backend, int variant) {
    switch (backend) {
    case kKS: return persistent_fn<kKS>(variant);
    case kPlain: return persistent_fn<kPlain>(variant);
    case kFancy: return persistent_fn<kFancy>(variant);
    default: return persistent_fn<kByte>(variant);
    }
}


int expand_grid(int backend, int mode, int small, int n_cu) {
    return blocks_per_cu(pick_expand(backend, mode, small)) * n_cu;
}
int persistent_max_bpc(int backend, int variant) { return blocks_per_cu(pick_persistent(backend, variant)); }

hipError_t launch_init_roots(int backend, const DevTables& dt, Board* roots, u32 n,
                             unsigned long long* sums, hipStream_t s) {
    if (!n) return hipSuccess;
    const dim3 g((n + kWG - 1) / kWG), blk(kWG);
    switch (backend) {
    case kKS: hipLaunchKernelGGL(k_init_roots<kKS>, g, blk, 0, s, dt.t, roots, n, sums); break;
    case kPlain: hipLaunchKernelGGL(k_init_roots<kPlain>, g, blk, 0, s, dt.t, roots, n, sums); break;
    case kFancy: hipLaunchKernelGGL(k_init_roots<kFancy>, g, blk, 0, s, dt.t, roots, n, sums); break;
    default: hipLaunchKernelGGL(k_init_roots<kByte>, g, blk, 0, s, dt.t, roots, n, sums); break;
    }
    return hipGetLastError();
}

hipError_t launch_expand(int backend, int mode, int small, const DevTables& dt, const LevelArgs& a, int grid,
                         hipStream_t s) {
    if (!a.n_in) return hipSuccess;
    void* f = pick_expand(backend, mode, small);
    LevelArgs args = a;
    const SliderTables* t = dt.t;
    void* kargs[] = {(void*)&t, (void*)&args};
    return hipLaunchKernel(f, dim3(grid), dim3(kWG), kargs, 0, s);
}

hipError_t launch_persistent(int backend, int variant, const DevTables& dt, const PersistArgs& a, int grid,
                             hipStream_t s) {
    void* f = pick_persistent(backend, variant);
    PersistArgs args = a;
    const SliderTables* t = dt.t;
    void* kargs[] = {(void*)&t, (void*)&args};
    return hipLaunchCooperativeKernel(f, dim3(grid), dim3(kWG), kargs, 0, s);
}

hipError_t launch_range_count(const Board* b, u32 n, unsigned long long* out, hipStream_t s) {
    if (!n) return hipSuccess;
    u32 g = (n + kWG * 8 - 1) / (kWG * 8);
    if (g > 2048) g = 2048;
    hipLaunchKernelGGL(k_range_count, dim3(g), dim3(kWG), 0, s, b, n, out);
    return hipGetLastError();
}

}  // namespace pf
