// level_dev.h -- the LDS-tiled level routine shared by the table-free
// kernels (kernels.hip: host-driven levels and the cooperative batch;
// stream.hip: the stream-ordered batch) and their tile shapes.
#pragma once
#include <hip/hip_runtime.h>

#include "backends_dev.h"
#include "kernels.h"
#include "wave.h"

namespace pf {

// Every level kernel runs 256-thread workgroups, several per CU, and the
// magic backends, byte lookup included, read their tables through the
// caches.  The byte backend's stream-ordered kernel has one more form, its
// 146 KiB of tables staged in the workgroup's LDS (below).
constexpr int kWG = 256;


// The byte backend's LDS form: the 148 KB lookup blob and the 4 KB of magic
// entries are copied into LDS once per workgroup (the workgroup is persistent,
// so the copy is amortised over every tile it processes).  That leaves room
// for one workgroup per CU, so it is 768 threads wide (three waves per SIMD)
// and its tile is 64 parents x 1920 moves, the largest that fits beside the
// tables.  On the 256 x perft 5 batch it runs 6.3e11 leaves/s against 1.15e12
// for the cache-reading form with five 256-thread workgroups per CU: every
// lookup is two dependent reads, and three waves per SIMD cannot hide them.
constexpr int kByteLdsWG = 768;
constexpr int kByteLdsTP = 64;
constexpr int kByteLdsCap = 1920;
constexpr size_t kByteBlob = offsetof(SliderTables, byte_b) + sizeof(u64) * 1428 - offsetof(SliderTables, byte_tab);
constexpr size_t kByteEntries = sizeof(MagicEntry) * 128;
constexpr size_t kByteLds = kByteBlob + kByteEntries;
static_assert(kByteBlob % 16 == 0 && offsetof(SliderTables, byte_tab) % 16 == 0, "16-byte staging");

__device__ __forceinline__ ByteBackend stage_byte_tables(const SliderTables* t, unsigned char* dyn) {
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
static_assert(sizeof(TileLds<kByteLdsTP, kByteLdsCap, kByteLdsWG>) + kByteLds + 128 <= 160 * 1024,
              "byte backend's LDS form: tables plus tile state must fit one CU's LDS");

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
        const u32 incl = wave_incl_scan(cnt);
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

// Tile shapes of the persistent kernel that the tuner chooses between:
// variant 0 = 256 parents x 8192-entry move window, 1 = 128 x 4096,
// 2 = 256 x 4096.
template <int B, int V>
struct PCfg {
    static constexpr int TP = (V == 1) ? 128 : 256;
    static constexpr int CAP = (V == 0) ? 8192 : 4096;
};

// Tile shape of the per-level kernels (k_expand, k_stream_level): 256 x 4096
// (variant 2), whose smaller LDS footprint lets more workgroups share a CU
// than the 256 x 8192 shape.
template <int B>
struct LevelV {
    static constexpr int v = 2;
};
// Waves per SIMD the per-level kernels are built for (1 = no constraint).
// The leaf kernel is built for five, so five workgroups share a CU; the magic
// backends' fits 96 VGPRs as it is, Kogge-Stone's spills a few bytes per lane
// (outside its make+count loop).  C2 leaves/s of the stream-ordered form,
// 256 x 8192 at three per CU -> 256 x 4096 at four -> at five: plain 1.05e12
// -> 1.18e12 -> 1.30e12, fancy 1.08e12 -> 1.20e12 -> 1.31e12 (their table
// reads from L2 want the extra waves), Kogge-Stone 1.470e12 -> 1.479e12 ->
// 1.500e12.  The write kernels would spill more and stay at four.
template <int B, int MODE>
struct LevelWaves {
    static constexpr int v = (MODE == kLeaf) ? 5 : 1;
};
// Workgroups of one of these kernels that fit a CU (a kernel with dynamic LDS
// has it enabled first).
static inline int blocks_per_cu(void* f, int wg = kWG, size_t dyn = 0) {
    if (dyn && hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess)
        (void)hipGetLastError();  // the occupancy query below then reports what fits
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, wg, dyn) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();  // a failed query must not surface as a later call's error
        per_cu = 1;
    }
    return per_cu;
}

}  // namespace pf
