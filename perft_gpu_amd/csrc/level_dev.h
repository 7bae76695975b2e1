// level_dev.h -- the LDS-tiled level routine shared by the table-free
// kernels (kernels.hip: host-driven levels and the cooperative batch;
// stream.hip: the stream-ordered batch) and their tile shapes.
#pragma once
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
// Occupancy of one of these kernels (the byte backend's dynamic LDS is
// enabled first).
static inline size_t dyn_bytes(int backend) { return backend == kByte ? Cfg<kByte>::dyn : 0; }
static inline int wg_size(int backend) { return backend == kByte ? kByteWG : kWG; }

static inline int blocks_per_cu(void* f, int backend) {
    if (backend == kByte) hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_bytes(backend));
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, wg_size(backend), dyn_bytes(backend)) != hipSuccess ||
        per_cu < 1) {
        (void)hipGetLastError();  // a failed query must not surface as a later call's error
        per_cu = 1;
    }
    return per_cu;
}

}  // namespace pf
