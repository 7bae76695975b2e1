// stream.hip -- the batch as a sequence of launches on one stream, one per
// level (PF_MODE_STREAM, and PF_MODE_AUTO's table-free single pass).
//
// The levels are sized and placed on the device exactly as in k_persistent
// (the totals and level offsets live in PersistCtl), so nothing returns to
// the host between levels, and the launch boundary replaces the grid barrier
// (the reference syncs the device after every level, perft_bb.h:1071-1135).
// Every workgroup of a level launch derives the same placement from words
// the previous launch finished writing.
//
// The level kernel is built in kStreamVariants forms -- two tile shapes times
// three waves-per-SIMD targets -- for the leaf mode and in the first five of
// them for the write mode, and the autotuner picks the leaf and the write
// form per backend (the reference fixes its launch shape at compile time,
// perft_bb.h:327-329, switches.h:14-21).
//
// The write kernel's 256 x 4096 form built for five waves per SIMD (96
// VGPRs, over 100 B of spills per lane) is not built: compiled from the
// current make-move and count code with the current compiler, it gave counts
// that differ from the oracle's on a perft 4 sample, and differently from run
// to run.  Its other forms, the leaf kernel's five-wave forms, and the same
// form built with only one of the two latest changes to that code (the
// flag decoding of make_move, the branch-free diagonal masks) all match.  No
// cause in the source was found: the LDS tile is fenced by barriers and every
// spill slot lies inside the declared private segment.  The form was never a
// tuned choice: the write level spills more than it gains at that target.
#include <hip/hip_runtime.h>

#include "level_dev.h"

namespace pf {

// Variant v = shape * 3 + waves: shape 0 is the 256 x 8192 tile (PCfg
// variant 0), shape 1 the 256 x 4096 tile (PCfg variant 2); waves 0, 1, 2 ask
// the compiler for no target, four and five resident waves per SIMD.
template <int SH>
struct StreamShape {
    static constexpr int v = SH ? 2 : 0;
};
template <int WV>
struct StreamWaves {
    static constexpr int v = WV == 2 ? 5 : WV == 1 ? 4 : 1;
};

template <int B>
__global__ __launch_bounds__(kWG) void k_stream_roots(const SliderTables* tabs, PersistArgs p) {
    const auto sl = SlOf<B>::make(tabs);
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long c = 0;
    if (i < p.n_roots) {
        const Board b = p.roots[i];
        c = count_moves(b, sl);
        p.roots[i] = with_count(b, (u32)c);
        if (p.depth == 1) p.sums[p.root_ids[i]] = c;
    }
    c = wave_sum64(c);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&p.ctl->totals[1], c);
}

// One level for one workgroup of the launch.  `make_sl` builds the slider
// backend once the workgroup knows it has work (the byte backend's LDS form
// copies its tables into LDS there).
template <int MODE, int TP, int CAP, int WG, class MakeSl>
__device__ __forceinline__ void stream_level(const MakeSl& make_sl, const PersistArgs& p, int L,
                                             TileLds<TP, CAP, WG>& s) {
    PersistCtl* ctl = p.ctl;
    // an earlier level did not fit: nothing below it is counted (one read per
    // workgroup, so that the whole workgroup leaves together)
    if (threadIdx.x == 0) s.ok = ctl->err;
    __syncthreads();
    if (s.ok) return;
    const Board* cur = p.roots;
    const u32* cur_r = p.root_ids;
    u64 n = p.n_roots, off = 0;
    if (L > 0) {
        n = ctl->totals[L];
        off = ctl->lvl_off[L];
        cur = reinterpret_cast<const Board*>(p.arena + off);
        cur_r = reinterpret_cast<const u32*>(p.arena + off + ((n * sizeof(Board) + 255) & ~255ULL));
        off += ((n * sizeof(Board) + 255) & ~255ULL) + ((n * sizeof(u32) + 255) & ~255ULL);
    }
    LevelArgs a;
    a.in = cur;
    a.in_root = cur_r;
    a.n_in = (u32)n;
    a.tile_ctr = &ctl->tile_ctr[L];
    a.err = &ctl->err;
    a.out_count = &ctl->out_count[L];
    a.tile_parents = level_tile_parents(n, gridDim.x, TP);
    // workgroups past the level's tile count leave at once (block 0 stays:
    // it places the next level); the others drain the tile queue
    const u64 ntiles = (n + a.tile_parents - 1) / a.tile_parents;
    if (MODE == kLeaf) {
        if (blockIdx.x >= ntiles) return;
        const auto sl = make_sl();
        a.out = nullptr;
        a.out_root = nullptr;
        a.out_cap = 0;
        a.root_sums = p.sums;
        a.total = nullptr;
        expand_level(sl, a, kLeaf, s);
        return;
    }
    const u64 nn = ctl->totals[L + 1];
    const u64 bbytes = (nn * sizeof(Board) + 255) & ~255ULL, rbytes = (nn * sizeof(u32) + 255) & ~255ULL;
    if (nn > 0xFFFFFFFFull || off + bbytes + rbytes > p.arena_bytes) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&ctl->err, 2u);
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->lvl_off[L + 1] = off;
    if (blockIdx.x > 0 && blockIdx.x >= ntiles) return;
    const auto sl = make_sl();
    a.out = reinterpret_cast<Board*>(p.arena + off);
    a.out_root = reinterpret_cast<u32*>(p.arena + off + bbytes);
    a.out_cap = nn;
    a.root_sums = nullptr;
    a.total = &ctl->totals[L + 2];
    expand_level(sl, a, kWrite, s);
}

template <int B, int MODE, int SH, int WV>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(StreamWaves<WV>::v)))
void k_stream_level(const SliderTables* tabs, PersistArgs p, int L) {
    using Shape = PCfg<B, StreamShape<SH>::v>;
    __shared__ TileLds<Shape::TP, Shape::CAP, kWG> s;
    stream_level<MODE>([&] { return SlOf<B>::make(tabs); }, p, L, s);
}

// The byte backend's LDS form (level_dev.h): one 768-thread workgroup per CU,
// tables staged in its LDS.
template <int MODE>
__global__ __launch_bounds__(kByteLdsWG) void k_stream_level_byte_lds(const SliderTables* tabs, PersistArgs p, int L) {
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    __shared__ TileLds<kByteLdsTP, kByteLdsCap, kByteLdsWG> s;
    stream_level<MODE>([&] { return stage_byte_tables(tabs, dyn_lds); }, p, L, s);
}

// ---- launch plumbing --------------------------------------------------------
template <int B, int MODE>
static void* level_fn(int v) {
    switch (v) {
    case 0: return reinterpret_cast<void*>(&k_stream_level<B, MODE, 0, 0>);
    case 1: return reinterpret_cast<void*>(&k_stream_level<B, MODE, 0, 1>);
    case 2: return reinterpret_cast<void*>(&k_stream_level<B, MODE, 0, 2>);
    case 3: return reinterpret_cast<void*>(&k_stream_level<B, MODE, 1, 0>);
    case 4: return reinterpret_cast<void*>(&k_stream_level<B, MODE, 1, 1>);
    default:
        if constexpr (MODE == kWrite) return reinterpret_cast<void*>(&k_stream_level<B, MODE, 1, 1>);
        else return reinterpret_cast<void*>(&k_stream_level<B, MODE, 1, 2>);
    }
}
template <int B>
static void* level_fn_mode(int mode, int v) {
    return mode == kWrite ? level_fn<B, kWrite>(v) : level_fn<B, kLeaf>(v);
}
static void* pick_stream_level(int backend, int mode, int v) {
    if (backend == kByte && v == kStreamByteLds)
        return mode == kWrite ? reinterpret_cast<void*>(&k_stream_level_byte_lds<kWrite>)
                              : reinterpret_cast<void*>(&k_stream_level_byte_lds<kLeaf>);
    switch (backend) {
    case kKS: return level_fn_mode<kKS>(mode, v);
    case kPlain: return level_fn_mode<kPlain>(mode, v);
    case kFancy: return level_fn_mode<kFancy>(mode, v);
    default: return level_fn_mode<kByte>(mode, v);
    }
}
static void* pick_stream_roots(int backend) {
    switch (backend) {
    case kKS: return reinterpret_cast<void*>(&k_stream_roots<kKS>);
    case kPlain: return reinterpret_cast<void*>(&k_stream_roots<kPlain>);
    case kFancy: return reinterpret_cast<void*>(&k_stream_roots<kFancy>);
    default: return reinterpret_cast<void*>(&k_stream_roots<kByte>);
    }
}

// Workgroup size and dynamic LDS of a level kernel form.
static int stream_wg(int backend, int variant) {
    return backend == kByte && variant == kStreamByteLds ? kByteLdsWG : kWG;
}
static size_t stream_dyn(int backend, int variant) {
    return backend == kByte && variant == kStreamByteLds ? kByteLds : 0;
}

int stream_grid(int backend, int mode, int variant, int n_cu) {
    return blocks_per_cu(pick_stream_level(backend, mode, variant), stream_wg(backend, variant),
                         stream_dyn(backend, variant)) * n_cu;
}

hipError_t launch_stream(int backend, const DevTables& dt, const PersistArgs& a, const int grid[2],
                         const int variant[2], hipStream_t s) {
    PersistArgs args = a;
    const SliderTables* t = dt.t;
    {
        void* kargs[] = {(void*)&t, (void*)&args};
        const hipError_t e = hipLaunchKernel(pick_stream_roots(backend), dim3((a.n_roots + kWG - 1) / kWG), dim3(kWG),
                                             kargs, 0, s);
        if (e != hipSuccess) return e;
    }
    for (int L = 0; L <= a.depth - 2; L++) {
        const int mode = (L == a.depth - 2) ? kLeaf : kWrite;
        void* kargs[] = {(void*)&t, (void*)&args, (void*)&L};
        // (a form with dynamic LDS had the attribute set when its grid was sized)
        void* f = pick_stream_level(backend, mode, variant[mode]);
        const size_t dyn = stream_dyn(backend, variant[mode]);
        const hipError_t e = hipLaunchKernel(f, dim3(grid[mode]), dim3(stream_wg(backend, variant[mode])), kargs,
                                             dyn, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace pf
