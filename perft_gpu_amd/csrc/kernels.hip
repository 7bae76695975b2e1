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

#include "backends_dev.h"
#include "kernels.h"

namespace pf {

constexpr int kWG = 256;

// The byte-lookup backend fills nearly all of a CU's LDS with its tables, so a
// CU holds a single one of its workgroups.  That workgroup is larger (12
// waves, three per SIMD) so that other waves hide each wave's table-lookup
// latency; the other backends run several 256-thread workgroups per CU.
// Measured on C2 (256 positions, perft 5): 256 threads 3.7e11 leaves/s,
// 512 5.4e11, 768 6.2e11, 1024 5.9e11 (all with the 64 x 1920 tile below).
constexpr int kByteWG = 768;
template <int B>
struct WgOf {
    static constexpr int v = (B == kByte) ? kByteWG : kWG;
};

__device__ __forceinline__ u32 wave_incl_scan(u32 v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const u32 o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Byte backend: the 148 KB lookup blob and the 4 KB of magic entries are copied
// into LDS once per workgroup; the workgroup is persistent, so the copy is
// amortised over every tile it processes.
constexpr size_t kByteBlob = offsetof(SliderTables, byte_b) + sizeof(u64) * 1428 - offsetof(SliderTables, byte_tab);
constexpr size_t kByteEntries = sizeof(MagicEntry) * 128;
constexpr size_t kByteLds = kByteBlob + kByteEntries;
static_assert(kByteBlob % 16 == 0 && offsetof(SliderTables, byte_tab) % 16 == 0, "16-byte staging");

template <int B>
__device__ __forceinline__ typename SlOf<B>::T stage_backend(const SliderTables* t, unsigned char* dyn) {
    if constexpr (B == kByte) {
        const uint4* se = reinterpret_cast<const uint4*>(t->fancy_b);
        uint4* de = reinterpret_cast<uint4*>(dyn);
        for (u32 i = threadIdx.x; i < kByteEntries / 16; i += blockDim.x) de[i] = se[i];
        const uint4* sb = reinterpret_cast<const uint4*>(t->byte_tab);
        uint4* db = reinterpret_cast<uint4*>(dyn + kByteEntries);
        for (u32 i = threadIdx.x; i < kByteBlob / 16; i += blockDim.x) db[i] = sb[i];
        __syncthreads();
        const MagicEntry* e = reinterpret_cast<const MagicEntry*>(dyn);
        const unsigned char* base = dyn + kByteEntries;
        ByteBackend bb;
        bb.eb = e;
        bb.er = e + 64;
        bb.bt = base;
        bb.ur = reinterpret_cast<const u64*>(base + (offsetof(SliderTables, byte_r) - offsetof(SliderTables, byte_tab)));
        bb.ub = reinterpret_cast<const u64*>(base + (offsetof(SliderTables, byte_b) - offsetof(SliderTables, byte_tab)));
        return bb;
    } else {
        (void)dyn;
        return SlOf<B>::make(t);
    }
}

// Emitter used by generate(): stores the moves of one parent whose index in
// the tile's move sequence falls inside the current LDS window.
struct LdsEmit {
    u32* list;
    int pos;  // index of the next move relative to the window start
    int cap;
    u32 tag;  // parent slot << 16
    __device__ __forceinline__ void operator()(u32 m) {
        if (pos >= 0 && pos < cap) list[pos] = m | tag;
        pos++;
    }
};

template <int TP, int CAP, int WG>
struct TileLds {
    Board b[TP];
    u32 root[TP];
    u32 sum[TP];
    u32 mv[CAP];
    u32 wtot[WG / 64];
    unsigned long long red[WG / 64];
    u32 tile, base, ok;
};

// Process every tile of one level (tiles come from a.tile_ctr).
template <class SL, int TP, int CAP, int WG>
__device__ __forceinline__ void expand_level(const SL& sl, const LevelArgs& a, const int mode,
                                             TileLds<TP, CAP, WG>& s) {
    static_assert(TP <= WG && (TP & (TP - 1)) == 0, "one parent per thread, power of two");
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const u32 tp = (a.tile_parents && a.tile_parents < (u32)TP) ? a.tile_parents : (u32)TP;
    const u32 ntiles = (a.n_in + tp - 1) / tp;
    for (;;) {
        if (tid == 0) s.tile = atomicAdd(a.tile_ctr, 1u);
        __syncthreads();
        const u32 tile = s.tile;
        if (tile >= ntiles) break;
        const u32 p0 = tile * tp;
        const u32 np = min(tp, a.n_in - p0);
        u32 cnt = 0;
        if (tid < TP) {
            if ((u32)tid < np) {
                const Board b = a.in[p0 + tid];
                s.b[tid] = b;
                s.root[tid] = a.in_root[p0 + tid];
                cnt = (u32)meta_count(b);
            }
            s.sum[tid] = 0;
        }
        const u32 incl = wave_incl_scan(cnt, lane);
        if (lane == 63) s.wtot[wid] = incl;
        __syncthreads();
        u32 woff = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < WG / 64; w++) {
            const u32 t = s.wtot[w];
            woff += (w < wid) ? t : 0;
            tot += t;
        }
        const u32 pre = woff + incl - cnt;

        for (u32 w0 = 0; w0 < tot; w0 += CAP) {
            if (cnt && pre < w0 + CAP && pre + cnt > w0) {
                const Board b = s.b[tid];
                PosInfo pi;
                analyse(b, sl, pi);
                LdsEmit e{s.mv, (int)pre - (int)w0, CAP, (u32)tid << 16};
                generate(b, sl, pi, e);
            }
            const u32 m = min((u32)CAP, tot - w0);
            if (mode == kWrite && tid == 0) {
                const u32 base = atomicAdd(a.out_count, m);
                const bool ok = (u64)base + m <= a.out_cap;
                s.base = base;
                s.ok = ok;
                if (!ok) atomicOr(a.err, 1u);
            }
            __syncthreads();
            const bool wr = (mode == kWrite) && s.ok;
            const u32 base = s.base;
            for (u32 j = tid; j < m; j += WG) {
                const u32 e = s.mv[j];
                const u32 par = (e >> 16) & (TP - 1);
                const Board c = make_move(s.b[par], e & 0xFFFFu);
                const u32 cc = count_moves(c, sl);
                atomicAdd(&s.sum[par], cc);
                if (wr) {
                    a.out[base + j] = with_count(c, cc);
                    a.out_root[base + j] = s.root[par];
                }
            }
            __syncthreads();
        }

        const u32 ps = ((u32)tid < np) ? s.sum[tid] : 0u;
        const u32 r0 = s.root[0];
        const int uni = __syncthreads_and(((u32)tid >= np) || (s.root[tid < TP ? tid : 0] == r0));
        if (uni || a.total) {
            const unsigned long long v = wave_sum64((unsigned long long)ps);
            if (lane == 0) s.red[wid] = v;
            __syncthreads();
            if (tid == 0) {
                unsigned long long t = 0;
#pragma unroll
                for (int w = 0; w < WG / 64; w++) t += s.red[w];
                if (t && a.total) atomicAdd(a.total, t);
                if (t && uni && a.root_sums) atomicAdd(&a.root_sums[r0], t);
            }
        }
        if (!uni && ps && a.root_sums) atomicAdd(&a.root_sums[s.root[tid]], (unsigned long long)ps);
        __syncthreads();
    }
}

template <int B>
struct Cfg {
    static constexpr int TP = 256;
    static constexpr int CAP = 8192;
    static constexpr size_t dyn = 0;
};
// Byte backend tile: 64 parents and a 1920-entry move window, the largest
// window that fits beside the staged tables, so that a tile's moves (about 30
// per parent) mostly fit one window and each parent is generated once.
template <>
struct Cfg<kByte> {
    static constexpr int TP = 64;
    static constexpr int CAP = 1920;
    static constexpr size_t dyn = kByteLds;
};
static_assert(sizeof(TileLds<Cfg<kByte>::TP, Cfg<kByte>::CAP, kByteWG>) + kByteLds + 128 <= 160 * 1024,
              "byte backend: tables plus tile state must fit one CU's LDS");

// Tile shapes of the persistent kernel that the tuner chooses between:
// variant 0 = 256 parents x 8192-entry move window, 1 = 128 x 4096,
// 2 = 256 x 4096.  The byte-lookup backend keeps its LDS-staged shape.
template <int B, int V>
struct PCfg {
    static constexpr int TP = (V == 1) ? 128 : 256;
    static constexpr int CAP = (V == 0) ? 8192 : 4096;
};
template <int V>
struct PCfg<kByte, V> {
    static constexpr int TP = Cfg<kByte>::TP;
    static constexpr int CAP = Cfg<kByte>::CAP;
};

// Tile shape of the per-level kernels (k_expand, k_stream_level): 256 x 4096
// (variant 2), whose smaller LDS footprint lets more workgroups share a CU
// than the 256 x 8192 shape.  The byte backend keeps its own shape.
template <int B>
struct LevelV {
    static constexpr int v = (B == kByte) ? 0 : 2;
};
// Waves per SIMD the per-level kernels are built for (1 = no constraint).
// The leaf kernel is built for five, so five workgroups share a CU; the magic
// backends' fits 96 VGPRs as it is, Kogge-Stone's spills 28 B per lane
// (outside its make+count loop).  C2 leaves/s of the stream-ordered form,
// 256 x 8192 at three per CU -> 256 x 4096 at four -> at five: plain 1.05e12
// -> 1.18e12 -> 1.30e12, fancy 1.08e12 -> 1.20e12 -> 1.31e12 (their table
// reads from L2 want the extra waves), Kogge-Stone 1.470e12 -> 1.479e12 ->
// 1.500e12.  The write kernels would spill more and stay at four.
template <int B, int MODE>
struct LevelWaves {
    static constexpr int v = (B != kByte && MODE == kLeaf) ? 5 : 1;
};

// The host-driven level loop has both shapes: the wide one above for large
// levels, and for small ones (few parents per workgroup, latency-bound) the
// 256 x 8192 tile without a waves-per-SIMD target, which does not spill.
template <int B, int MODE, int V>
__global__ __launch_bounds__(WgOf<B>::v)
__attribute__((amdgpu_waves_per_eu(V == LevelV<B>::v ? LevelWaves<B, MODE>::v : 1)))
void k_expand(const SliderTables* tabs, LevelArgs a) {
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    __shared__ TileLds<PCfg<B, V>::TP, PCfg<B, V>::CAP, WgOf<B>::v> s;
    const auto sl = stage_backend<B>(tabs, dyn_lds);
    expand_level(sl, a, MODE, s);
}

template <int B>
__global__ __launch_bounds__(kWG) void k_init_roots(const SliderTables* t, Board* roots, u32 n,
                                                    unsigned long long* sums) {
    constexpr int G = (B == kByte) ? kFancy : B;  // roots are few: no LDS staging
    const typename SlOf<G>::T sl = SlOf<G>::make(t);
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
__global__ __launch_bounds__(WgOf<B>::v) void k_persistent(const SliderTables* tabs, PersistArgs p) {
    namespace cg = cooperative_groups;
    cg::grid_group grid = cg::this_grid();
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    __shared__ TileLds<PCfg<B, V>::TP, PCfg<B, V>::CAP, WgOf<B>::v> s;
    const auto sl = stage_backend<B>(tabs, dyn_lds);
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

// The same batch as a sequence of launches on one stream, one per level: the
// levels are sized and placed on the device exactly as in k_persistent (the
// totals and level offsets live in PersistCtl), so nothing returns to the
// host between levels, and the launch boundary replaces the grid barrier.
// Every workgroup of a level launch derives the same placement from words
// the previous launch finished writing.
template <int B>
__global__ __launch_bounds__(WgOf<B>::v) void k_stream_roots(const SliderTables* tabs, PersistArgs p) {
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    const auto sl = stage_backend<B>(tabs, dyn_lds);
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

template <int B, int MODE>
__global__ __launch_bounds__(WgOf<B>::v) __attribute__((amdgpu_waves_per_eu(LevelWaves<B, MODE>::v)))
void k_stream_level(const SliderTables* tabs, PersistArgs p, int L) {
    extern __shared__ __align__(16) unsigned char dyn_lds[];
    __shared__ TileLds<PCfg<B, LevelV<B>::v>::TP, PCfg<B, LevelV<B>::v>::CAP, WgOf<B>::v> s;
    PersistCtl* ctl = p.ctl;
    // an earlier level did not fit: nothing below it is counted (one read per
    // workgroup, so that the whole workgroup leaves together)
    if (threadIdx.x == 0) s.ok = ctl->err;
    __syncthreads();
    if (s.ok) return;
    const auto sl = stage_backend<B>(tabs, dyn_lds);
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
    a.tile_parents = level_tile_parents(n, gridDim.x, PCfg<B, LevelV<B>::v>::TP);
    if (MODE == kLeaf) {
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
    a.out = reinterpret_cast<Board*>(p.arena + off);
    a.out_root = reinterpret_cast<u32*>(p.arena + off + bbytes);
    a.out_cap = nn;
    a.root_sums = nullptr;
    a.total = &ctl->totals[L + 2];
    expand_level(sl, a, kWrite, s);
}

// ---- launch plumbing --------------------------------------------------------
template <int B, int MODE>
static void* expand_fn(int small) {
    if (B == kByte || small) return reinterpret_cast<void*>(&k_expand<B, MODE, 0>);
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
    default: return reinterpret_cast<void*>(&k_persistent<kByte, 0>);
    }
}

static size_t dyn_bytes(int backend) { return backend == kByte ? Cfg<kByte>::dyn : 0; }
static int wg_size(int backend) { return backend == kByte ? kByteWG : kWG; }

static int blocks_per_cu(void* f, int backend) {
    if (backend == kByte) hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_bytes(backend));
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, wg_size(backend), dyn_bytes(backend)) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return per_cu;
}

int expand_grid(int backend, int mode, int small, int n_cu) {
    return blocks_per_cu(pick_expand(backend, mode, small), backend) * n_cu;
}
int persistent_max_bpc(int backend, int variant) { return blocks_per_cu(pick_persistent(backend, variant), backend); }

hipError_t launch_init_roots(int backend, const DevTables& dt, Board* roots, u32 n,
                             unsigned long long* sums, hipStream_t s) {
    if (!n) return hipSuccess;
    const dim3 g((n + kWG - 1) / kWG), blk(kWG);
    switch (backend) {
    case kKS: hipLaunchKernelGGL(k_init_roots<kKS>, g, blk, 0, s, dt.t, roots, n, sums); break;
    case kPlain: hipLaunchKernelGGL(k_init_roots<kPlain>, g, blk, 0, s, dt.t, roots, n, sums); break;
    default: hipLaunchKernelGGL(k_init_roots<kFancy>, g, blk, 0, s, dt.t, roots, n, sums); break;
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
    return hipLaunchKernel(f, dim3(grid), dim3(wg_size(backend)), kargs, dyn_bytes(backend), s);
}

hipError_t launch_persistent(int backend, int variant, const DevTables& dt, const PersistArgs& a, int grid,
                             hipStream_t s) {
    void* f = pick_persistent(backend, variant);
    PersistArgs args = a;
    const SliderTables* t = dt.t;
    void* kargs[] = {(void*)&t, (void*)&args};
    return hipLaunchCooperativeKernel(f, dim3(grid), dim3(wg_size(backend)), kargs, (unsigned)dyn_bytes(backend), s);
}

template <int B>
static void* stream_level_fn(int mode) {
    return mode ? reinterpret_cast<void*>(&k_stream_level<B, kWrite>) : reinterpret_cast<void*>(&k_stream_level<B, kLeaf>);
}
static void* pick_stream_level(int backend, int mode) {
    switch (backend) {
    case kKS: return stream_level_fn<kKS>(mode);
    case kPlain: return stream_level_fn<kPlain>(mode);
    case kFancy: return stream_level_fn<kFancy>(mode);
    default: return stream_level_fn<kByte>(mode);
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

int stream_grid(int backend, int mode, int n_cu) { return blocks_per_cu(pick_stream_level(backend, mode), backend) * n_cu; }

hipError_t launch_stream(int backend, const DevTables& dt, const PersistArgs& a, const int grid[2], hipStream_t s) {
    PersistArgs args = a;
    const SliderTables* t = dt.t;
    const unsigned dyn = (unsigned)dyn_bytes(backend);
    const int wg = wg_size(backend);
    if (backend == kByte) hipFuncSetAttribute(pick_stream_roots(backend), hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    {
        void* kargs[] = {(void*)&t, (void*)&args};
        const hipError_t e = hipLaunchKernel(pick_stream_roots(backend), dim3((a.n_roots + wg - 1) / wg), dim3(wg),
                                             kargs, dyn, s);
        if (e != hipSuccess) return e;
    }
    for (int L = 0; L <= a.depth - 2; L++) {
        const int mode = (L == a.depth - 2) ? kLeaf : kWrite;
        void* kargs[] = {(void*)&t, (void*)&args, (void*)&L};
        const hipError_t e = hipLaunchKernel(pick_stream_level(backend, mode), dim3(grid[mode]), dim3(wg), kargs, dyn, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_range_count(const Board* b, u32 n, unsigned long long* out, hipStream_t s) {
    if (!n) return hipSuccess;
    u32 g = (n + kWG * 8 - 1) / (kWG * 8);
    if (g > 2048) g = 2048;
    hipLaunchKernelGGL(k_range_count, dim3(g), dim3(kWG), 0, s, b, n, out);
    return hipGetLastError();
}

}  // namespace pf
