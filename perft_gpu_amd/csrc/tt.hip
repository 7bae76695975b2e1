// tt.hip -- breadth-first perft with the transposition table.
//
// Same tile structure as kernels.hip (parents in LDS, their legal moves in an
// LDS list, one (parent, move) pair per lane), plus:
//   * every child carries two incremental Zobrist keys (tt.h) and the index
//     of its parent;
//   * expand-and-write: each child of remaining depth >= kMinTTDepth probes
//     its bucket; a hit adds the stored count to the parent's accumulator and
//     the child is not materialised (replaces
//     makemove_and_count_moves_single_level_hash128b, perft_bb.h:664-786);
//     a miss claims its slot in the level's duplicate table: the first child
//     of a position is compacted with a wave ballot and written, later ones
//     only join the level's duplicate list
//   * leaf: each parent's accumulator receives perft(parent, 2)
//   * upsweep: a finished level adds its counts, and its duplicates their
//     claimants' counts, into the parents' accumulators and stores the nodes
//     in the table (calcPerftNFromPerftNminus1_hash128b, perft_bb.h:854-961)
#include <hip/hip_runtime.h>

#include "tt_kernels.h"
#include "wave.h"

namespace pf {

constexpr int kTWG = 256;
// Parents per tile: 112 parents have about 3,600 moves (32.5 per node on
// average), so a tile's moves mostly fit one 4096-entry window; 128 had
// about 4,200 and spilled a few dozen into a second window almost every tile.
constexpr int kTTP = kTTTileParents;
constexpr int kTCAP = 4096; // LDS move-list window
// Move-list windows a tile can need (a position has at most 218 legal moves).
constexpr int kTTMaxWin = 8;
static_assert(kTTP * 218 <= kTTMaxWin * kTCAP, "a tile's moves fit kTTMaxWin windows");
// The write kernels wait on random table and duplicate-claim traffic most of
// the time: ask for at least this many resident waves per SIMD (register
// budget 512 / waves) so that more of that latency is hidden.
constexpr int kTTWaves = 4;
// The level driver's leaf kernel is built for five waves per SIMD (it spills
// 40 B per lane for Kogge-Stone): C3 perft 8 0.0879 -> 0.0868 s and C5
// 2.189 -> 2.162 s against four waves without spills.
constexpr int kTTLeafWaves = 5;
// The serving step's mixed launch (write and leaf tiles at once) is built for
// five too: 96 VGPRs, 96 B per lane of spills (none in the leaf loop), five
// workgroups per CU.  On the 6,018-request serving stream submitted at once:
// 7.60e12 leaves/s for the write and leaf launches one after the other,
// 7.96e12 mixed at four waves, 8.04-8.12e12 at five, 8.36-8.40e12 once a tile
// is taken with one atomic.
constexpr int kTTMixedWaves = 5;

// Adds each lane's v to its *p with one global atomic per run of consecutive
// lanes that hold the same address.  A level's children are written in their
// parents' order, so the 64 items of an upsweep wave usually belong to two or
// three parents; the run's head lane adds the run's total.  Every lane of the
// wave must call it; a lane with p == nullptr adds nothing.
__device__ __forceinline__ void wave_run_add(unsigned long long* p, unsigned long long v) {
    const int lane = threadIdx.x & 63;
    const unsigned long long a = reinterpret_cast<unsigned long long>(p);
    const unsigned long long prev = __shfl_up(a, 1, 64);
    const bool head = lane == 0 || prev != a;
    const unsigned long long heads = __ballot(head);
    const u32 rid = (u32)__popcll(heads & ((2ULL << lane) - 1));  // runs that start at or below this lane
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {  // suffix sums within each run
        const unsigned long long ov = __shfl_down(v, d, 64);
        const u32 orid = __shfl_down(rid, d, 64);
        if (lane + d < 64 && orid == rid) v += ov;
    }
    if (head && p && v) atomicAdd(p, v);
}

__device__ __forceinline__ bool tt_probe(const uint4* tt, u64 mask, u64 k1, u64 k2, int depth, u64* val) {
    const uint4* bkt = tt + 2 * tt_bucket(k1, depth, mask);
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const uint4 e = bkt[s];
        const u64 w0 = (u64)e.x | ((u64)e.y << 32), w1 = (u64)e.z | ((u64)e.w << 32);
        if ((w0 ^ w1) == k2 && (int)(w1 >> 56) == depth) {
            *val = w1 & kCountMask;
            return true;
        }
    }
    return false;
}

__device__ __forceinline__ void tt_store(uint4* tt, u64 mask, u64 k1, u64 k2, int depth, u64 val) {
    if (val > kCountMask) return;
    uint4* bkt = tt + 2 * tt_bucket(k1, depth, mask);
    const uint4 e0 = bkt[0];
    const u64 o1 = (u64)e0.z | ((u64)e0.w << 32);
    const int od = (int)(o1 >> 56);
    const u64 w1 = ((u64)depth << 56) | val, w0 = k2 ^ w1;
    const uint4 ne = make_uint4((u32)w0, (u32)(w0 >> 32), (u32)w1, (u32)(w1 >> 32));
    if (depth >= od) bkt[0] = ne;  // depth-preferred slot
    else bkt[1] = ne;              // always-replace slot
}

// Slot of a key in a duplicate table of n slots (n < 2^32): multiply-shift
// range reduction of the high half of the first key.
__device__ __forceinline__ u64 dup_slot(u64 k1, u64 n) { return ((k1 >> 32) * n) >> 32; }

// Outcomes of a made child in a write level (bits 0-1 of its code byte; bits
// 2-6 hold the probe distance of its duplicate-table slot).
enum ChildCode : u32 { kClaimed = 0, kDup = 1, kKeepNoSlot = 2, kSkip = 3 };

__device__ __forceinline__ u64 probe_slot(u64 k1, u64 n, u32 dist) {
    const u64 s = dup_slot(k1, n) + dist;
    return s >= n ? s - n : s;
}

// A made child claims the first empty slot of its probe sequence.  Returns
// kClaimed when it claimed slot probe_slot(k1, n, *dist) (the first child of
// its position in this level: it is written and publishes its index there),
// kDup when an earlier child with its keys holds that slot, kKeepNoSlot when
// the probe sequence is full or the key is 0 (written, no duplicate
// detection).
// A slot goes from 0 to a key once in a level's life, so a non-zero key read
// at device scope is final: a duplicate (most children at deep levels) is
// recognised by a load, and only an empty slot needs the compare-and-swap.
__device__ __forceinline__ u32 dup_claim(DupEntry* t, u64 n, u64 k1, u64 k2, u32* dist) {
    if (!k2) return kKeepNoSlot;
    u64 s = dup_slot(k1, n);
    for (u32 p = 0; p < (u32)kDupProbes; p++) {
        unsigned long long cur = __hip_atomic_load(&t[s].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == 0) {
            cur = atomicCAS(&t[s].key, 0ull, (unsigned long long)k2);
            if (cur == 0) {
                *dist = p;
                return kClaimed;
            }
        }
        if (cur == k2) {
            *dist = p;
            return kDup;
        }
        s = (s + 1 == n) ? 0 : s + 1;
    }
    return kKeepNoSlot;
}

static_assert(kDupProbes <= 32, "a probe distance fits in 5 bits of a child's code");

struct TTLds {
    Board b[kTTP];
    u32 root[kTTP];
    u32 sum[kTTP];
    u64 k1[kTTP], k2[kTTP];
    unsigned long long hit[kTTP];
    u32 mv[kTCAP];
    unsigned char code[kTCAP];  // write levels: ChildCode | probe distance << 2 per window entry
    u32 wtot[kTWG / 64];
    unsigned long long red[kTWG / 64];
    u32 tile;
    u32 nprobe, nhit;
    u32 nkeep, ndup;  // window totals from the claim pass
    u32 kcur, dcur;   // window-local cursors of the write pass
    u32 base, dbase;  // the window's ranges in the level and in the duplicate list
    u32 ok;
    u32 nwin;                    // write tiles: windows so far, and each one's
    u32 wr_base[kTTMaxWin];      // written children [wr_base, wr_base + wr_n)
    u32 wr_n[kTTMaxWin];
};

struct LdsEmitT {
    u32* list;
    int pos;
    int cap;
    u32 tag;
    __device__ __forceinline__ void operator()(u32 m) {
        if (pos >= 0 && pos < cap) list[pos] = m | tag;
        pos++;
    }
};

// The m children of one move-list window of a write level.  Every thread of
// the block calls it; it ends with the window's writes issued.
// Two passes, so that the level's and the duplicate list's cursors -- one
// address each, shared by every workgroup, whose atomics serialise -- move
// once per window rather than once per wave and 64 children:
//  1. make, keys, table probe, duplicate claim: each child's outcome and
//     probe distance go to s.code, the window's counts to s.nkeep / s.ndup;
//  2. one thread reserves both ranges; every child to be written or listed
//     as a duplicate is made again with its keys and placed at a window-local
//     cursor inside the reserved range.
template <class SL>
__device__ __forceinline__ void write_window(const SL& sl, const TTArgs& a, const ZTables* zt, TTLds& s,
                                             const u32 p0, const u32 m, const int lane, const int wid) {
    const u64 below_mask = (1ULL << lane) - 1;
    u32 wkeep = 0, wdup = 0, wprobe = 0, whit = 0;  // wave-uniform sums
    for (u32 jb = (u32)wid * 64; jb < m; jb += kTWG) {
        const u32 j = jb + lane;
        u32 code = kSkip;
        bool probed = false;
        if (j < m) {
            const u32 e = s.mv[j];
            const u32 par = (e >> 16) & 0xFFu, mvv = e & 0xFFFFu;
            const Board pb = s.b[par];
            Board c = make_move(pb, mvv);
            if (a.ep_always) c = ep_on_every_double_push(c, mvv);
            u64 ck1 = 0, ck2 = 0;
            if (a.tt) {
                const int black = (int)(s.root[par] >> 31);
                ck1 = zobrist_child(zt->k[0], s.k1[par], pb, mvv, black, c);
                ck2 = zobrist_child(zt->k[1], s.k2[par], pb, mvv, black, c);
            }
            probed = a.tt && a.child_depth >= kMinTTDepth;
            u64 v = 0;
            if (probed && tt_probe(a.tt, a.tt_mask, ck1, ck2, a.child_depth, &v)) {
                atomicAdd(&s.hit[par], (unsigned long long)v);
            } else if (a.dup) {
                u32 dist = 0;
                code = dup_claim(a.dup, a.dup_n, ck1, ck2, &dist);
                code |= dist << 2;
            } else {
                code = kKeepNoSlot;
            }
            s.code[j] = (unsigned char)code;
        }
        const u32 res = code & 3u;
        wkeep += (u32)__popcll(__ballot(res == kClaimed || res == kKeepNoSlot));
        wdup += (u32)__popcll(__ballot(res == kDup));
        wprobe += (u32)__popcll(__ballot(probed));
        whit += (u32)__popcll(__ballot(probed && res == kSkip));
    }
    if (lane == 0) {
        if (wkeep) atomicAdd(&s.nkeep, wkeep);
        if (wdup) atomicAdd(&s.ndup, wdup);
        if (wprobe) {
            atomicAdd(&s.nprobe, wprobe);
            atomicAdd(&s.nhit, whit);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const u32 nk = s.nkeep, nd = s.ndup;
        const u32 base = nk ? atomicAdd(a.out_count, nk) : 0u;
        const u32 dbase = nd ? atomicAdd(a.dup_count, nd) : 0u;
        const bool ok = (u64)base + nk <= a.out_cap && (u64)dbase + nd <= a.out_cap;
        if (!ok) atomicOr(a.err, 1u);
        s.base = base;
        s.dbase = dbase;
        s.ok = ok;
        s.wr_base[s.nwin] = base;  // nwin < kTTMaxWin: a tile has at most that many windows
        s.wr_n[s.nwin] = ok ? nk : 0u;
        s.nwin++;
        s.kcur = 0;
        s.dcur = 0;
        s.nkeep = 0;
        s.ndup = 0;
    }
    __syncthreads();
    if (!s.ok) return;  // uniform: the level overflowed its exact size
    for (u32 jb = (u32)wid * 64; jb < m; jb += kTWG) {
        const u32 j = jb + lane;
        u32 code = kSkip, par = 0, mvv = 0;
        if (j < m) {
            code = s.code[j];
            const u32 e = s.mv[j];
            par = (e >> 16) & 0xFFu;
            mvv = e & 0xFFFFu;
        }
        const u32 res = code & 3u;
        const bool keep = res == kClaimed || res == kKeepNoSlot, isdup = res == kDup;
        Board c = {};
        u64 ck1 = 0, ck2 = 0;
        u32 cc = 0;
        if (keep || isdup) {
            const Board pb = s.b[par];
            c = make_move(pb, mvv);
            if (a.ep_always) c = ep_on_every_double_push(c, mvv);
            if (a.tt) {
                const int black = (int)(s.root[par] >> 31);
                ck1 = zobrist_child(zt->k[0], s.k1[par], pb, mvv, black, c);
                if (keep) ck2 = zobrist_child(zt->k[1], s.k2[par], pb, mvv, black, c);
            }
        }
        if (keep) {
            cc = count_moves(c, sl);
            atomicAdd(&s.sum[par], cc);
        }
        const u64 kbal = __ballot(keep);
        const u32 nk = (u32)__popcll(kbal);
        u32 lb = 0;
        if (lane == 0 && nk) lb = atomicAdd(&s.kcur, nk);
        lb = __shfl(lb, 0, 64);
        const u64 first = (u64)s.base + lb;  // the wave's written children are [first, first + nk)
        const u64 pos = first + (u32)__popcll(kbal & below_mask);
        if (keep) {
            a.out[pos] = with_count(c, cc);
            a.out_root[pos] = s.root[par] ^ 0x80000000u;  // child's side to move
            if (a.tt) {
                a.out_k1[pos] = ck1;
                a.out_k2[pos] = ck2;
            }
            a.out_parent[pos] = a.in_base + p0 + par;
            a.out_acc[pos] = 0;
            if (res == kClaimed) a.dup[probe_slot(ck1, a.dup_n, code >> 2)].val = pos;  // publish the claimant
        }
        const u64 dbal = __ballot(isdup);
        const u32 nd = (u32)__popcll(dbal);
        if (nd) {
            u32 dl = 0;
            if (lane == 0) dl = atomicAdd(&s.dcur, nd);
            dl = __shfl(dl, 0, 64);
            if (isdup) {
                const u64 dpos = (u64)s.dbase + dl + (u32)__popcll(dbal & below_mask);
                a.dup_par[dpos] = a.in_base + p0 + par;
                a.dup_slot[dpos] = (u32)probe_slot(ck1, a.dup_n, code >> 2);
            }
        }
        if (a.out_bsum && nk) {
            // the wave's children are consecutive: at most two blocks
            const u64 b0 = first >> kBsumShift;
            const u64 part = keep ? (((pos >> kBsumShift) == b0) ? (u64)cc : (u64)cc << 32) : 0;
            const u64 t = wave_sum64(part);
            if (lane == 0) {
                if ((u32)t) atomicAdd(&a.out_bsum[b0], (u32)t);
                if (t >> 32) atomicAdd(&a.out_bsum[b0 + 1], (u32)(t >> 32));
            }
        }
    }
}

// One tile (kTTP parents) of one level.  Every thread of the block calls it.
template <class SL>
__device__ void tt_tile(const SL& sl, const TTArgs& a, const u32 tile, const ZTables* zt, TTLds& s) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    {
        if (tid == 0) {
            s.nprobe = 0;
            s.nhit = 0;
            s.nkeep = 0;
            s.ndup = 0;
            s.nwin = 0;
        }
        const u32 p0 = tile * kTTP;
        const u32 np = min((u32)kTTP, a.n_in - p0);
        u32 cnt = 0;
        if (tid < kTTP) {
            if ((u32)tid < np) {
                const Board b = a.in[p0 + tid];
                s.b[tid] = b;
                s.root[tid] = a.in_root[p0 + tid];
                if (a.tt) {
                    s.k1[tid] = a.in_k1[p0 + tid];
                    s.k2[tid] = a.in_k2[p0 + tid];
                }
                cnt = (u32)meta_count(b);
            }
            s.sum[tid] = 0;
            s.hit[tid] = 0;
        }
        const u32 incl = wave_incl_scan(cnt);
        if (lane == 63) s.wtot[wid] = incl;
        __syncthreads();
        u32 woff = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < kTWG / 64; w++) {
            const u32 t = s.wtot[w];
            woff += (w < wid) ? t : 0;
            tot += t;
        }
        const u32 pre = woff + incl - cnt;
        for (u32 w0 = 0; w0 < tot; w0 += kTCAP) {
            if (cnt && pre < w0 + kTCAP && pre + cnt > w0) {
                const Board b = s.b[tid];
                PosInfo pi;
                analyse(b, sl, pi);
                LdsEmitT e{s.mv, (int)pre - (int)w0, kTCAP, (u32)tid << 16};
                generate(b, sl, pi, e);
            }
            __syncthreads();
            const u32 m = min((u32)kTCAP, tot - w0);
            if (a.mode == kLeaf) {
                // each wave walks 64 consecutive entries at a time (uniform loop)
                for (u32 jb = (u32)wid * 64; jb < m; jb += kTWG) {
                    const u32 j = jb + lane;
                    if (j < m) {
                        const u32 e = s.mv[j];
                        const u32 par = (e >> 16) & 0xFFu;
                        const u32 cc = count_moves(make_move(s.b[par], e & 0xFFFFu), sl);
                        atomicAdd(&s.sum[par], cc);
                    }
                }
            } else {
                write_window(sl, a, zt, s, p0, m, lane, wid);
            }
            __syncthreads();
        }
        // accumulators: hits (write mode) or perft(parent, 2) (leaf mode)
        unsigned long long tsum = 0;
        if ((u32)tid < np) {
            const unsigned long long add = (a.mode == kWrite) ? s.hit[tid] : (unsigned long long)s.sum[tid];
            a.in_acc[p0 + tid] += add;
            tsum = s.sum[tid];
        }
        if (a.total) {
            const unsigned long long v = wave_sum64(tsum);
            if (lane == 0) s.red[wid] = v;
        }
        __syncthreads();
        if (tid == 0) {
            if (a.total) {
                unsigned long long t = 0;
                for (int w = 0; w < kTWG / 64; w++) t += s.red[w];
                if (t) atomicAdd(a.total, t);
            }
            if (a.stats && s.nprobe) {
                atomicAdd(&a.stats[0], (unsigned long long)s.nprobe);
                atomicAdd(&a.stats[1], (unsigned long long)s.nhit);
            }
        }
        __syncthreads();
    }
}

template <class SL>
__device__ void expand_tt_level(const SL& sl, const TTArgs& a, const ZTables* zt, TTLds& s) {
    const u32 ntiles = (a.n_in + kTTP - 1) / kTTP;
    for (;;) {
        if (threadIdx.x == 0) s.tile = atomicAdd(a.tile_ctr, 1u);
        __syncthreads();
        const u32 tile = s.tile;
        __syncthreads();
        if (tile >= ntiles) break;
        tt_tile(sl, a, tile, zt, s);
    }
}

// One kernel per mode: with the mode known at compile time the leaf kernel
// carries none of the write pass, and is built for kTTLeafWaves waves per SIMD.
template <int B, int M>
__global__ __launch_bounds__(kTWG) __attribute__((amdgpu_waves_per_eu(M == kLeaf ? kTTLeafWaves : kTTWaves)))
void k_expand_tt(const SliderTables* tabs, const ZTables* zt, TTArgs a) {
    __shared__ TTLds s;
    const auto sl = SlOf<B>::make(tabs);
    a.mode = M;
    expand_tt_level(sl, a, zt, s);
}

// Roots: keys were computed on the host; probe at the full depth, count moves.
template <int B>
__global__ __launch_bounds__(256) void k_roots_tt(const SliderTables* tabs, Board* roots, const u64* k1,
                                                  const u64* k2, u32 n, int depth, const uint4* tt, u64 mask,
                                                  unsigned long long* acc, unsigned long long* total,
                                                  unsigned long long* stats) {
    const auto sl = SlOf<B>::make(tabs);
    const u32 i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const Board b = roots[i];
    u64 v = 0;
    if (tt && depth >= kMinTTDepth) {
        atomicAdd(&stats[0], 1ULL);
        if (tt_probe(tt, mask, k1[i], k2[i], depth, &v)) {
            atomicAdd(&stats[1], 1ULL);
            acc[i] = v;
            roots[i] = with_count(b, 0);  // answered: nothing to expand
            return;
        }
    }
    const u32 c = count_moves(b, sl);
    roots[i] = with_count(b, c);
    acc[i] = 0;
    atomicAdd(total, (unsigned long long)c);
}

// Nodes of an upsweep segment, and all its items (nodes, then duplicates):
// from the host, or from the outcome of a write that ran in the same step;
// none when that write overflowed, so that no count short of it is added up
// or stored in the table.
__device__ __forceinline__ u32 up_nodes(const UpSeg& u) {
    if (u.dev && u.dev->err) return 0;  // the write below it overflowed: nothing goes up or into the table
    return u.dev_counts ? u.dev->out_count : u.n;
}
__device__ __forceinline__ u32 up_items(const UpSeg& u) {
    if (u.dev && u.dev->err) return 0;
    return u.dev_counts ? u.dev->out_count + u.dev->ndup : u.n + u.ndup;
}

// Item k of a finished level's upsweep (calcPerftNFromPerftNminus1_hash128b,
// perft_bb.h:854-961) whose n nodes come first: k < n is node k, which adds
// its exact count to its parent and is stored in the table; n <= k < n + ndup
// is duplicate k - n, which adds the count of the node that claimed its slot.
// The addition is returned in (*dst, *v) for wave_run_add.  Returns 1 for a
// store.
__device__ __forceinline__ u32 upsweep_item(const UpSeg& u, u32 n, u32 k, uint4* tt, u64 mask,
                                            unsigned long long** dst, unsigned long long* v) {
    if (k < n) {
        *v = u.acc[k];
        *dst = u.parent_acc ? &u.parent_acc[u.parent[k]] : nullptr;
        if (tt && u.depth >= kMinTTDepth) {
            tt_store(tt, mask, u.k1[k], u.k2[k], u.depth, *v);
            return 1;
        }
        return 0;
    }
    const u32 d = k - n;
    if (u.parent_acc) {
        const u32 c = (u32)u.dup[u.dup_slot[d]].val;
        *v = u.acc[c];
        *dst = &u.parent_acc[u.dup_par[d]];
    }
    return 0;
}

__global__ __launch_bounds__(256) void k_upsweep(UpSeg u, uint4* tt, u64 mask, unsigned long long* stores) {
    const u32 i = blockIdx.x * 256 + threadIdx.x;
    unsigned long long st = 0, v = 0;
    unsigned long long* dst = nullptr;
    if (i < u.n + u.ndup) st = upsweep_item(u, u.n, i, tt, mask, &dst, &v);
    wave_run_add(dst, v);
    if (stores) {
        st = wave_sum64(st);
        if ((threadIdx.x & 63) == 0 && st) atomicAdd(stores, st);
    }
}

// ---- serving step: many jobs' segments in one launch ---------------------------
// A segment's arguments, read from LDS, as wave-uniform values: each dword
// goes through readfirstlane, so the pointers and sizes sit in scalar
// registers as a kernel argument's would, instead of taking some forty
// vector registers (and spilling) in every lane.
__device__ __forceinline__ TTArgs uniform_args(const TTArgs& src) {
    static_assert(sizeof(TTArgs) % 4 == 0, "TTArgs is copied dword by dword");
    u32 w[sizeof(TTArgs) / 4];
    __builtin_memcpy(w, &src, sizeof(w));
#pragma unroll
    for (int i = 0; i < (int)(sizeof(TTArgs) / 4); i++) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
    TTArgs a;
    __builtin_memcpy(&a, w, sizeof(a));
    return a;
}

// Tiles of the step's expand segments form a queue (StepTable::tile_prefix);
// a block copies its segment's arguments to LDS and runs the tile.  Segments
// of different jobs touch disjoint memory except the lockless table.
// W: the waves per SIMD it is built for; M = kWrite runs the table's write
// segments [0, nseg_write) and M = kLeaf its leaf segments [nseg_write,
// nseg), each without the other mode's code (as in k_expand_tt).
template <int B, int W, int M>
__global__ __launch_bounds__(kTWG) __attribute__((amdgpu_waves_per_eu(W))) void k_step(const SliderTables* tabs, const ZTables* zt, StepTable* t) {
    __shared__ TTLds s;
    __shared__ TTArgs sa;
    __shared__ u32 s_tile, s_seg;
    const auto sl = SlOf<B>::make(tabs);
    const u32 s_lo = (M == kLeaf) ? t->nseg_write : 0u, s_hi = (M == kWrite) ? t->nseg_write : t->nseg;
    const u32 first = t->tile_prefix[s_lo], total = t->tile_prefix[s_hi];
    if (M == kWrite && t->fuse_leaf) return;  // the fused launch runs this level
    if (blockIdx.x >= total - first) return;  // more workgroups than tiles: the others drain the queue
    u32* const ctr = (M == kLeaf) ? &t->tile_ctr2 : &t->tile_ctr;
    u32 cur = 0xFFFFFFFFu;
    for (;;) {
        if (threadIdx.x == 0) {
            const u32 g = first + atomicAdd(ctr, 1u);
            u32 seg = s_lo;
            if (g < total) {
                u32 lo = s_lo, hi = s_hi;  // largest seg with tile_prefix[seg] <= g
                while (hi - lo > 1) {
                    const u32 mid = (lo + hi) >> 1;
                    if (t->tile_prefix[mid] <= g) lo = mid; else hi = mid;
                }
                seg = lo;
            }
            s_tile = g;
            s_seg = seg;
        }
        __syncthreads();
        const u32 g = s_tile, seg = s_seg;
        if (g >= total) break;
        if (seg != cur) {
            if (threadIdx.x == 0) sa = t->seg[seg];
            cur = seg;
        }
        __syncthreads();
        TTArgs a = uniform_args(sa);
        a.mode = M;
        tt_tile(sl, a, g - t->tile_prefix[seg], zt, s);
    }
}

// The planned pass's last write level and its leaf pass in one launch (its
// segment 0, a write level whose children have remaining depth 2), when the
// level's planner chose it (StepTable::fuse_leaf: enough tiles for every
// workgroup; otherwise this launch leaves at once and the write and leaf
// launches run the level).  A workgroup writes one tile's children and then
// runs the leaf pass over the children it kept, at once, while they are still
// in the caches.  Workgroups are in different phases at any time, so the
// write tiles' waits on table and duplicate claims overlap other workgroups'
// leaf VALU work, where a write launch followed by a leaf launch ran them one
// after the other: C5's first 2048 positions at perft 6 0.934 -> 0.865 s, C3
// perft 8 0.080 -> 0.076 s.  The kept children of a tile are the windows'
// reserved ranges (TTLds::wr_base/wr_n), whose accumulators the write just
// zeroed; the leaf tiles add perft(child, 2) to them.
template <int B, int W>
__global__ __launch_bounds__(kTWG) __attribute__((amdgpu_waves_per_eu(W))) void k_step_write_leaf(const SliderTables* tabs, const ZTables* zt, StepTable* t) {
    __shared__ TTLds s;
    __shared__ TTArgs sa;
    __shared__ u32 s_tile;
    const auto sl = SlOf<B>::make(tabs);
    const u32 total = t->tile_prefix[1];
    if (!t->fuse_leaf) return;  // the level was written by the write launch; the leaf launch follows
    if (blockIdx.x >= total) return;  // more workgroups than tiles: the others drain the queue
    if (threadIdx.x == 0) sa = t->seg[0];
    __syncthreads();
    TTArgs a = uniform_args(sa);
    a.mode = kWrite;
    TTArgs l = {};  // the leaf tiles' arguments: parents are the written children
    l.mode = kLeaf;
    l.child_depth = a.child_depth - 1;
    l.err = a.err;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&t->tile_ctr, 1u);
        __syncthreads();
        const u32 g = s_tile;
        if (g >= total) break;
        tt_tile(sl, a, g, zt, s);  // ends with a barrier: s.nwin and the ranges are final
        const u32 nwin = s.nwin;
        __syncthreads();  // every wave has nwin before a leaf tile reuses the tile state
        for (u32 w = 0; w < nwin; w++) {
            const u32 rb = s.wr_base[w], rn = s.wr_n[w];
            for (u32 c0 = 0; c0 < rn; c0 += kTTP) {
                l.in = a.out + rb + c0;
                l.in_root = a.out_root + rb + c0;
                l.in_acc = a.out_acc + rb + c0;
                l.n_in = min((u32)kTTP, rn - c0);
                tt_tile(sl, l, 0, zt, s);
            }
        }
    }
}

// The serving step's write and leaf segments in one launch, overlapped:
// workgroups [0, mix_write_wgs) start on the write tiles, whose waves mostly
// wait on table and duplicate-table accesses, the others on the leaf tiles,
// which keep the VALU busy; a workgroup whose queue is empty moves to the
// other one, and leaves when that is empty too.  The queues are the write
// segments [0, nseg_write) with tile_ctr and the leaf segments [nseg_write,
// nseg) with tile_ctr2; their bounds are read once, before the loop, and the
// tile prefix is copied to LDS, so that taking a tile is one atomic and a
// search in LDS (small steps are latency-bound).
template <int B, int W, bool F>
__global__ __launch_bounds__(kTWG) __attribute__((amdgpu_waves_per_eu(W))) void k_step_mixed(const SliderTables* tabs, const ZTables* zt, StepTable* t) {
    __shared__ TTLds s;
    __shared__ TTArgs sa;
    __shared__ u32 s_tile, s_seg;
    __shared__ u32 pre[kMaxSegs + 1];  // the tile prefix, searched for every tile
    const auto sl = SlOf<B>::make(tabs);
    const u32 nw = t->nseg_write, ns = t->nseg;
    const u32 wt = t->tile_prefix[nw], lt = t->tile_prefix[ns];  // write tiles [0, wt), leaf tiles [wt, lt)
    if (blockIdx.x >= lt) return;  // more workgroups than tiles: the others drain the queues
    const u32 tp = F ? t->mix_tp : (u32)kTTP;
    for (u32 k = threadIdx.x; k <= ns; k += kTWG) pre[k] = t->tile_prefix[k];
    __syncthreads();
    bool leaf = blockIdx.x >= t->mix_write_wgs;  // thread 0's queue
    bool moved = false;
    u32 cur = 0xFFFFFFFFu;
    for (;;) {
        if (threadIdx.x == 0) {
            u32 g = 0xFFFFFFFFu;
            for (;;) {
                const u32 first = leaf ? wt : 0u, total = leaf ? lt : wt;
                if (first < total) {
                    const u32 k = first + atomicAdd(leaf ? &t->tile_ctr2 : &t->tile_ctr, 1u);
                    if (k < total) {
                        g = k;
                        break;
                    }
                }
                if (moved) break;
                leaf = !leaf;
                moved = true;
            }
            u32 seg = 0;
            if (g != 0xFFFFFFFFu) {
                u32 lo = leaf ? nw : 0u, hi = leaf ? ns : nw;  // largest seg with pre[seg] <= g
                while (hi - lo > 1) {
                    const u32 mid = (lo + hi) >> 1;
                    if (pre[mid] <= g) lo = mid; else hi = mid;
                }
                seg = lo;
            }
            s_tile = g;
            s_seg = seg;
        }
        __syncthreads();
        const u32 g = s_tile, seg = s_seg;
        if (g == 0xFFFFFFFFu) break;
        if (seg != cur) {
            if (threadIdx.x == 0) sa = t->seg[seg];
            cur = seg;
        }
        __syncthreads();
        const TTArgs a = uniform_args(sa);  // the segment's own mode
        if (F) {  // tile g - pre[seg] of the segment: its tp parents as a one-tile range
            const u32 c0 = (g - pre[seg]) * tp;
            TTArgs f = a;
            f.in += c0;
            f.in_root += c0;
            f.in_acc += c0;
            f.in_base += c0;
            if (f.in_k1) {
                f.in_k1 += c0;
                f.in_k2 += c0;
            }
            f.n_in = min(tp, a.n_in - c0);
            tt_tile(sl, f, 0, zt, s);
        } else {
            tt_tile(sl, a, g - pre[seg], zt, s);
        }
    }
}

// Serving: the leaf passes over levels written by this step's write launch
// (StepTable::late_pos), so that a job's level of remaining depth 2 is
// written, counted and upswept in one step.  Each pass's parents are its write
// segment's children; their number is that write's outcome (none if it
// overflowed), so the tile prefix is built here, by the first wave.
template <int B, int W>
__global__ __launch_bounds__(kTWG) __attribute__((amdgpu_waves_per_eu(W))) void k_step_late(const SliderTables* tabs, const ZTables* zt, StepTable* t) {
    __shared__ TTLds s;
    __shared__ u32 pre[kMaxLate + 1];
    __shared__ u32 s_tile, s_seg, s_tp;
    const u32 nl = t->nlate;
    if (!nl) return;
    const auto sl = SlOf<B>::make(tabs);
    if (threadIdx.x < 64) {
        u32 n = 0;
        if (threadIdx.x < nl) {
            const TTArgs& w = t->seg[t->late_pos[threadIdx.x]];
            n = *w.err ? 0u : *w.out_count;
        }
        // parents per tile: the fewest that still give every tile a workgroup
        // of this launch (these levels are small and wait on the chain of one
        // tile's dependent steps), at most a full tile
        u32 tp = kTTP;
        for (u32 c = kFineMinParents; c < (u32)kTTP; c *= 2)
            if (__builtin_amdgcn_readlane(wave_incl_scan((n + c - 1) / c), 63) <= gridDim.x) {
                tp = c;
                break;
            }
        const u32 incl = wave_incl_scan((n + tp - 1) / tp);
        if (threadIdx.x < nl) pre[threadIdx.x + 1] = incl;
        if (threadIdx.x == 0) {
            pre[0] = 0;
            s_tp = tp;
        }
    }
    __syncthreads();
    const u32 total = pre[nl], tp = s_tp;
    if (blockIdx.x >= total) return;  // more workgroups than tiles: the others drain the queue
    u32 cur = 0xFFFFFFFFu, wn = 0;
    TTArgs l = {}, wl = {};
    for (;;) {
        if (threadIdx.x == 0) {
            const u32 g = atomicAdd(&t->tile_ctr3, 1u);
            u32 lo = 0, hi = nl;  // largest k with pre[k] <= g
            if (g < total)
                while (hi - lo > 1) {
                    const u32 mid = (lo + hi) >> 1;
                    if (pre[mid] <= g) lo = mid; else hi = mid;
                }
            s_tile = g;
            s_seg = lo;
        }
        __syncthreads();
        const u32 g = s_tile, k = s_seg;
        if (g >= total) break;
        if (k != cur) {  // the write's children as parents (uniform: from LDS and the table)
            const TTArgs& w = t->seg[t->late_pos[k]];
            wl = w;
            wn = *w.out_count;
            l.mode = kLeaf;
            l.child_depth = w.child_depth - 1;
            l.err = w.err;
            cur = k;
        }
        // tile g - pre[k] of the level: its tp parents as a one-tile range
        const u32 c0 = (g - pre[k]) * tp;
        l.in = wl.out + c0;
        l.in_root = wl.out_root + c0;
        l.in_acc = wl.out_acc + c0;
        l.n_in = min(tp, wn - c0);
        tt_tile(sl, l, 0, zt, s);
    }
}

// Upsweep round `round`: item i of the round's concatenated segments (nodes
// and duplicates) finds its segment by search.  Every workgroup first builds
// the round's item prefix in LDS, as a segment written in this same step has
// its counts only on the device.
__global__ __launch_bounds__(256) void k_up_step(const StepTable* __restrict__ t, uint4* tt, u64 mask, int round) {
    __shared__ u32 pre[kMaxUps + 1];
    __shared__ u32 wtot[4];
    const u32 s0 = t->up_round[round], s1 = t->up_round[round + 1], ns = s1 - s0;
    // pre[k] = items of segments [0, k): thread x sums segments 2x and 2x + 1
    const u32 k0 = 2 * threadIdx.x, k1 = k0 + 1;
    const u32 c0 = k0 < ns ? up_items(t->up[s0 + k0]) : 0u, c1 = k1 < ns ? up_items(t->up[s0 + k1]) : 0u;
    const u32 incl = wave_incl_scan(c0 + c1);
    if ((threadIdx.x & 63) == 63) wtot[threadIdx.x >> 6] = incl;
    __syncthreads();
    u32 woff = 0, last = 0;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        woff += (w < (int)(threadIdx.x >> 6)) ? wtot[w] : 0u;
        last += wtot[w];
    }
    const u32 excl = woff + incl - c0 - c1;
    if (k0 < ns) pre[k0] = excl;
    if (k1 < ns) pre[k1] = excl + c0;
    if (threadIdx.x == 0) pre[ns] = last;
    __syncthreads();
    if (blockIdx.x * 256 >= last) return;  // no item for this workgroup
    unsigned long long stores = 0;
    // the loop is uniform per wave (wave_run_add needs every lane)
    for (u32 w = blockIdx.x * 256 + (threadIdx.x & ~63u); w < last; w += gridDim.x * 256) {
        const u32 i = w + (threadIdx.x & 63);
        unsigned long long v = 0;
        unsigned long long* dst = nullptr;
        u32 lo = 0;
        if (i < last) {
            u32 hi = ns;  // largest segment with pre[lo] <= i
            while (hi - lo > 1) {
                const u32 mid = (lo + hi) >> 1;
                if (pre[mid] <= i) lo = mid; else hi = mid;
            }
        }
        // most waves lie in one segment: its arguments then come through
        // scalar loads, once per wave
        const u32 lo0 = __builtin_amdgcn_readfirstlane(lo);
        if (__all(i >= last || lo == lo0)) {
            if (i < last) {
                const UpSeg& u = t->up[s0 + lo0];
                stores += upsweep_item(u, up_nodes(u), i - pre[lo0], tt, mask, &dst, &v);
            }
        } else if (i < last) {
            const UpSeg& u = t->up[s0 + lo];
            stores += upsweep_item(u, up_nodes(u), i - pre[lo], tt, mask, &dst, &v);
        }
        wave_run_add(dst, v);
    }
    if (t->stores) {
        stores = wave_sum64(stores);
        if ((threadIdx.x & 63) == 0 && stores) atomicAdd(t->stores, stores);
    }
}

// Serving: the step table's host part, read by the GPU from coherent host
// memory -- only what the step uses (header, tile prefix, upsweep rounds and
// the used segments, upsweeps and roots), so that a small step moves about a
// kilobyte rather than the whole table, and no copy node sits between the
// host and the step's first kernel.  Copied as 4-byte words.
__global__ __launch_bounds__(256) void k_fetch_step(StepTable* __restrict__ d, const StepTable* __restrict__ h) {
    __shared__ u32 hdr[3];
    if (threadIdx.x == 0) {
        hdr[0] = h->nseg;
        hdr[1] = h->nup;
        hdr[2] = h->nroot;
    }
    __syncthreads();
    // (clamped to the table's capacities, so every word copied lies inside it)
    const u32 nseg = min(hdr[0], (u32)kMaxSegs), nup = min(hdr[1], (u32)kMaxUps), nroot = min(hdr[2], (u32)kMaxRoots);
    constexpr u32 kHdr = offsetof(StepTable, tile_prefix) / 4;
    const u32 o_tp = offsetof(StepTable, tile_prefix) / 4, n_tp = nseg + 1;
    const u32 o_ur = offsetof(StepTable, up_round) / 4, n_ur = kUpRounds + 1;
    const u32 o_sg = offsetof(StepTable, seg) / 4, n_sg = nseg * (u32)(sizeof(TTArgs) / 4);
    const u32 o_up = offsetof(StepTable, up) / 4, n_up = nup * (u32)(sizeof(UpSeg) / 4);
    const u32 o_rt = offsetof(StepTable, root) / 4, n_rt = nroot * (u32)(sizeof(RootSeg) / 4);
    const u32 total = kHdr + n_tp + n_ur + n_sg + n_up + n_rt;
    const u32* src = reinterpret_cast<const u32*>(h);
    u32* dst = reinterpret_cast<u32*>(d);
    for (u32 i = blockIdx.x * 256 + threadIdx.x; i < total; i += gridDim.x * 256) {
        u32 k = i, w;  // word k of the concatenated ranges -> its word in the table
        if (k < kHdr) w = k;
        else if ((k -= kHdr) < n_tp) w = o_tp + k;
        else if ((k -= n_tp) < n_ur) w = o_ur + k;
        else if ((k -= n_ur) < n_sg) w = o_sg + k;
        else if ((k -= n_sg) < n_up) w = o_up + k;
        else w = o_rt + (k - n_up);
        dst[w] = src[w];
    }
}
static_assert(sizeof(TTArgs) % 4 == 0 && sizeof(UpSeg) % 4 == 0 && sizeof(RootSeg) % 4 == 0 &&
                  offsetof(StepTable, tile_prefix) % 4 == 0,
              "the step table is fetched in 4-byte words");

// Serving: the step's outcomes back into the coherent host table (the used
// segment outcomes, root outcomes, the slot accumulators and the table
// counters); then, once all of them are visible to the host, the step's
// number into done, so that the host can wait for the step by watching it.
// One workgroup.
__global__ __launch_bounds__(256) void k_post_step(const StepTable* d, StepTable* h, u32 max_jobs) {
    const u32 nseg = min(d->nseg, (u32)kMaxSegs), nroot = min(d->nroot, (u32)kMaxRoots);
    max_jobs = min(max_jobs, (u32)kMaxRoots);
    for (u32 i = threadIdx.x; i < nseg; i += 256) h->out[i] = d->out[i];
    for (u32 i = threadIdx.x; i < nroot; i += 256) h->rout[i] = d->rout[i];
    for (u32 i = threadIdx.x; i < max_jobs; i += 256) h->acc[i] = d->acc[i];
    if (threadIdx.x < 2) h->stats[threadIdx.x] = d->stats[threadIdx.x];
    __threadfence_system();  // this thread's writes reach the host before the barrier ...
    __syncthreads();
    if (threadIdx.x == 0)    // ... and so before done
        __hip_atomic_store(&h->done, d->seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void k_post_flag(u32* flag, u32 v) {
    if (threadIdx.x == 0) __hip_atomic_store(flag, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

hipError_t launch_post_flag(u32* flag, u32 v, hipStream_t s) {
    hipLaunchKernelGGL(k_post_flag, dim3(1), dim3(64), 0, s, flag, v);
    return hipGetLastError();
}

hipError_t launch_fetch_step(StepTable* d, const StepTable* h, hipStream_t s) {
    hipLaunchKernelGGL(k_fetch_step, dim3(32), dim3(256), 0, s, d, h);
    return hipGetLastError();
}
hipError_t launch_post_step(const StepTable* d, StepTable* h, u32 max_jobs, hipStream_t s) {
    hipLaunchKernelGGL(k_post_step, dim3(1), dim3(256), 0, s, d, h, max_jobs);
    return hipGetLastError();
}

// Start of a step: the step's segment outcomes zeroed, new requests (table
// probe at full depth, move count) and empty duplicate tables for this step's
// write segments (a fill kernel: the step is a captured graph, which records
// no memset).
// `segs` is t->seg, only read here: the segments' arguments come through
// scalar loads.
template <int B>
__global__ __launch_bounds__(256) void k_pre_step(const SliderTables* tabs, StepTable* t,
                                                  const TTArgs* __restrict__ segs, const uint4* tt, u64 mask) {
    const auto sl = SlOf<B>::make(tabs);
    const u32 gi = blockIdx.x * 256 + threadIdx.x;
    if (gi < t->nseg) t->out[gi] = SegOut{};
    if (gi < t->nroot) {
        const RootSeg& r = t->root[gi];
        const Board b = r.b;
        *r.dst_w = r.word;
        if (tt) {
            *r.dst_k1 = r.k1;
            *r.dst_k2 = r.k2;
        }
        u64 v = 0;
        if (tt && r.depth >= kMinTTDepth && tt_probe(tt, mask, r.k1, r.k2, r.depth, &v)) {
            *r.acc = v;
            *r.dst_b = with_count(b, 0);  // answered: nothing below it is expanded
            t->rout[gi].hit = 1;
            t->rout[gi].count = 0;
        } else {
            const u32 c = count_moves(b, sl);
            *r.dst_b = with_count(b, c);
            *r.acc = (r.depth == 1) ? c : 0;
            t->rout[gi].hit = 0;
            t->rout[gi].count = c;
        }
    }
    const u64 stride = (u64)gridDim.x * 256;
    const u32 nw = t->nseg_write;  // leaf segments [nw, nseg) have no tables
    for (u32 si = 0; si < nw; si++) {
        const TTArgs& a = segs[si];
        if (a.out_bsum)
            for (u64 i = gi; i < a.bsum_n; i += stride) a.out_bsum[i] = 0;
        if (!a.dup) continue;
        for (u64 i = gi; i < a.dup_n; i += stride) a.dup[i] = DupEntry{0, 0};
    }
}

// Most resident workgroups per CU of the planned pass's step kernels.
int step_max_bpc(int backend, int leaf) {
    const void* f;
    if (leaf)
        f = backend == kKS ? (const void*)&k_step<kKS, kTTLeafWaves, kLeaf>
            : backend == kPlain ? (const void*)&k_step<kPlain, kTTLeafWaves, kLeaf>
            : backend == kFancy ? (const void*)&k_step<kFancy, kTTLeafWaves, kLeaf>
                                : (const void*)&k_step<kByte, kTTLeafWaves, kLeaf>;
    else
        f = backend == kKS ? (const void*)&k_step<kKS, kTTWaves, kWrite>
            : backend == kPlain ? (const void*)&k_step<kPlain, kTTWaves, kWrite>
            : backend == kFancy ? (const void*)&k_step<kFancy, kTTWaves, kWrite>
                                : (const void*)&k_step<kByte, kTTWaves, kWrite>;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kTWG, 0) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();  // a failed query must not surface as a later call's error
        per_cu = 1;
    }
    return per_cu;
}

// The serving step's mixed kernel: its most resident workgroups per CU x CUs.
int step_grid(int backend, int n_cu) {
    const void* f = backend == kKS ? (const void*)&k_step_mixed<kKS, kTTMixedWaves, false>
                    : backend == kPlain ? (const void*)&k_step_mixed<kPlain, kTTMixedWaves, false>
                    : backend == kFancy ? (const void*)&k_step_mixed<kFancy, kTTMixedWaves, false>
                                        : (const void*)&k_step_mixed<kByte, kTTMixedWaves, false>;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kTWG, 0) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    return per_cu * n_cu;
}

// One step: start (roots, fills), expand/leaf segments in one mixed launch,
// the late leaf passes (over levels the mixed launch wrote), up_rounds
// (<= kUpRounds) upsweep rounds.  up_grid * 256 must cover
// kMaxRoots and kMaxSegs (the roots and the outcomes are handled by the first
// threads of the start kernel).
// The step's start, mixed and (late) late launches for one backend and tiling.
template <int B, bool F>
static void launch_step_front(const SliderTables* tabs, const ZTables* zt, StepTable* t, uint4* tt, u64 mask, int grid,
                              int up_grid, int late, hipStream_t s) {
    hipLaunchKernelGGL(k_pre_step<B>, dim3(up_grid), dim3(256), 0, s, tabs, t, t->seg, tt, mask);
    hipLaunchKernelGGL((k_step_mixed<B, kTTMixedWaves, F>), dim3(grid), dim3(kTWG), 0, s, tabs, zt, t);
    if (late) hipLaunchKernelGGL((k_step_late<B, kTTLeafWaves>), dim3(grid), dim3(kTWG), 0, s, tabs, zt, t);
}

hipError_t launch_step(int backend, const SliderTables* tabs, const ZTables* zt, StepTable* t, uint4* tt, u64 mask,
                       int grid, int up_grid, int up_rounds, int fine, int late, hipStream_t s) {
    if (up_grid * 256 < kMaxRoots || up_grid * 256 < kMaxSegs) return hipErrorInvalidValue;
    if (up_rounds < 0 || up_rounds > kUpRounds) return hipErrorInvalidValue;
    using Front = void (*)(const SliderTables*, const ZTables*, StepTable*, uint4*, u64, int, int, int, hipStream_t);
    Front front;
    switch (backend) {
    case kKS: front = fine ? launch_step_front<kKS, true> : launch_step_front<kKS, false>; break;
    case kPlain: front = fine ? launch_step_front<kPlain, true> : launch_step_front<kPlain, false>; break;
    case kFancy: front = fine ? launch_step_front<kFancy, true> : launch_step_front<kFancy, false>; break;
    default: front = fine ? launch_step_front<kByte, true> : launch_step_front<kByte, false>; break;
    }
    front(tabs, zt, t, tt, mask, grid, up_grid, late, s);
    for (int r = 0; r < up_rounds; r++) hipLaunchKernelGGL(k_up_step, dim3(up_grid), dim3(256), 0, s, t, tt, mask, r);
    return hipGetLastError();
}

// ---- device-planned table pass -------------------------------------------------
// One-block planners between the launches of a captured pass (tt_kernels.h):
// they read the counts the previous launch left in the step table, place the
// next level in the pass's memory and write the next launch's segment.  A
// level that does not fit sets st.err; every later planner then leaves its
// launch empty, so the pass ends without an effect outside its memory.

// Takes `bytes` (rounded up to 256) of the pass's memory; on overflow sets err.
__device__ __forceinline__ char* plan_take(TTPlanCtl* c, u64 bytes) {
    const u64 b = (bytes + 255) & ~255ULL;
    if (c->st.err || c->st.peak + b > c->p.arena_bytes) {
        c->st.err = 2;
        return nullptr;
    }
    char* r = c->p.arena + c->st.peak;
    c->st.peak += b;
    return r;
}

// Empties the step table's launch descriptions (an empty step does nothing).
__device__ __forceinline__ void plan_empty(StepTable* t) {
    t->nseg = 0;
    t->nseg_write = 0;
    t->nup = 0;
    t->nroot = 0;
    t->tile_ctr = 0;
    t->tile_ctr2 = 0;
    t->tile_prefix[0] = 0;
    t->tile_prefix[1] = 0;
    t->fuse_leaf = 0;
    for (int r = 0; r <= kUpRounds; r++) t->up_round[r] = 0;
    t->up_prefix[0] = 0;
    t->out[0] = SegOut{};
}

// Level L of the pass as written by the step that just ran.
__device__ __forceinline__ void plan_record(TTPlanCtl* c, const StepTable* t, int L) {
    const SegOut& o = t->out[0];
    if (o.err) c->st.err = 2;
    TTLevelRec& lv = c->lv[L];
    lv.n = o.out_count;
    lv.ndup = o.ndup;
    lv.next = o.total;
    c->st.frontier += o.out_count;
    c->st.dups += o.ndup;
}

// Before the write of level L + 1 from level L.  At L = 0 it also sets up the
// pass: the parents (the caller's level, read only) with accumulators of the
// pass's own, and the children total from the parents' stored move counts.
__global__ __launch_bounds__(256) void k_plan_write(TTPlanCtl* c, StepTable* t, int L, uint4* tt, u64 mask,
                                                    unsigned long long* ttstats) {
    __shared__ unsigned long long red[4];
    if (L == 0) {
        const TTLevelRec& top = c->p.top;
        unsigned long long v = 0;
        for (u64 i = threadIdx.x; i < top.n; i += 256) v += (u64)meta_count(top.b[i]);
        v = wave_sum64(v);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (L == 0) {
        c->st = TTPlanStatus{};
        c->lv[0] = c->p.top;
        c->lv[0].next = red[0] + red[1] + red[2] + red[3];
        c->lv[0].acc = reinterpret_cast<unsigned long long*>(plan_take(c, c->p.top.n * 8));
    } else if (!c->st.err) {
        plan_record(c, t, L);
    }
    plan_empty(t);
    if (c->st.err) return;
    const TTLevelRec& cur = c->lv[L];
    const u64 cap = cur.next;
    if (cap > 0xFFFFFFFFull || cur.n > 0xFFFFFFFFull) {
        c->st.err = 2;
        return;
    }
    c->st.made += cap;
    TTLevelRec nx = {};
    nx.b = reinterpret_cast<Board*>(plan_take(c, cap * sizeof(Board)));
    nx.root = reinterpret_cast<u32*>(plan_take(c, cap * 4));
    nx.k1 = reinterpret_cast<u64*>(plan_take(c, cap * 8));
    nx.k2 = reinterpret_cast<u64*>(plan_take(c, cap * 8));
    nx.parent = reinterpret_cast<u32*>(plan_take(c, cap * 4));
    nx.acc = reinterpret_cast<unsigned long long*>(plan_take(c, cap * 8));
    const bool dd = c->p.dedup && c->p.ply + L + 1 >= kDupMinPly && cap >= 2;
    if (dd) {
        nx.dup_n = dup_table_slots(cap);
        nx.dup = reinterpret_cast<DupEntry*>(plan_take(c, nx.dup_n * sizeof(DupEntry)));
        nx.dpar = reinterpret_cast<u32*>(plan_take(c, cap * 4));
        nx.dslot = reinterpret_cast<u32*>(plan_take(c, cap * 4));
    }
    c->lv[L + 1] = nx;
    if (c->st.err) return;
    TTArgs a = {};
    a.in = cur.b;
    a.in_root = cur.root;
    a.in_k1 = cur.k1;
    a.in_k2 = cur.k2;
    a.n_in = (u32)cur.n;
    a.in_base = 0;
    a.in_acc = cur.acc;
    a.out = nx.b;
    a.out_root = nx.root;
    a.out_k1 = nx.k1;
    a.out_k2 = nx.k2;
    a.out_parent = nx.parent;
    a.out_acc = nx.acc;
    a.out_cap = cap;
    a.out_count = &t->out[0].out_count;
    a.total = &t->out[0].total;
    a.tt = tt;
    a.tt_mask = mask;
    a.child_depth = c->p.depth - (L + 1);
    a.mode = kWrite;
    a.ep_always = c->p.ep_always;
    a.err = &t->out[0].err;
    a.stats = ttstats;
    if (dd) {
        a.dup = nx.dup;
        a.dup_n = nx.dup_n;
        a.dup_count = &t->out[0].ndup;
        a.dup_par = nx.dpar;
        a.dup_slot = nx.dslot;
    }
    t->seg[0] = a;
    const u32 tiles = (u32)((cur.n + kTTP - 1) / kTTP);
    t->tile_prefix[1] = tiles;
    t->nseg = 1;
    t->nseg_write = 1;
    // the last write level (children of remaining depth 2) fuses its leaf pass
    // when every workgroup gets many tiles: a fused tile (its write, then the
    // leaf tiles of its kept children) is long, and a few of them would leave
    // most workgroups idle
    t->fuse_leaf = (a.child_depth == 2 && tiles >= c->p.fuse_min_tiles) ? 1u : 0u;
}

// Empties level L + 1's duplicate table and, at L = 0, the pass's accumulators.
__global__ __launch_bounds__(256) void k_plan_clear(TTPlanCtl* c, int L) {
    if (c->st.err) return;
    const u64 stride = (u64)gridDim.x * 256, i0 = (u64)blockIdx.x * 256 + threadIdx.x;
    const TTLevelRec& nx = c->lv[L + 1];
    for (u64 k = i0; k < nx.dup_n; k += stride) nx.dup[k] = DupEntry{0, 0};
    if (L == 0)
        for (u64 k = i0; k < c->lv[0].n; k += stride) c->lv[0].acc[k] = 0;
}

// After the last write level (level L, remaining depth 2): records it and,
// unless the write launch ran the leaf pass fused (StepTable::fuse_leaf),
// plans the leaf pass over it.
__global__ __launch_bounds__(64) void k_plan_leaf(TTPlanCtl* c, StepTable* t, int L,
                                                  unsigned long long* ttstats) {
    if (threadIdx.x != 0) return;
    const bool fused = t->fuse_leaf != 0;
    if (!c->st.err) plan_record(c, t, L);
    plan_empty(t);
    if (c->st.err) return;
    const TTLevelRec& cur = c->lv[L];
    c->st.made += cur.next;
    if (fused) return;  // the write launch ran the leaf pass: the leaf launch stays empty
    TTArgs a = {};
    a.in = cur.b;
    a.in_root = cur.root;
    a.in_k1 = cur.k1;
    a.in_k2 = cur.k2;
    a.n_in = (u32)cur.n;
    a.in_base = 0;
    a.in_acc = cur.acc;
    a.out_count = &t->out[0].out_count;
    a.child_depth = c->p.depth - (L + 1);
    a.mode = kLeaf;
    a.err = &t->out[0].err;
    a.stats = ttstats;
    t->seg[0] = a;
    t->tile_prefix[1] = (u32)((cur.n + kTTP - 1) / kTTP);
    t->nseg = 1;
}

// Before the upsweep of level L + 1 into level L.
__global__ __launch_bounds__(64) void k_plan_up(TTPlanCtl* c, StepTable* t, int L, unsigned long long* ttstats) {
    if (threadIdx.x != 0) return;
    plan_empty(t);
    if (c->st.err) return;
    const TTLevelRec& ch = c->lv[L + 1];
    UpSeg u = {};
    u.parent = ch.parent;
    u.k1 = ch.k1;
    u.k2 = ch.k2;
    u.acc = ch.acc;
    u.parent_acc = c->lv[L].acc;
    u.dup = ch.dup;
    u.dup_par = ch.dpar;
    u.dup_slot = ch.dslot;
    u.n = (u32)ch.n;
    u.ndup = (u32)ch.ndup;
    u.depth = c->p.depth - (L + 1);
    t->up[0] = u;
    t->nup = 1;
    for (int r = 1; r <= kUpRounds; r++) t->up_round[r] = 1;
    t->up_prefix[1] = u.n + u.ndup;
    t->stores = ttstats ? ttstats + 2 : nullptr;
}

// The pass's counts into the caller's accumulators, if nothing overflowed.
__global__ __launch_bounds__(256) void k_plan_commit(TTPlanCtl* c) {
    if (c->st.err) return;
    const u64 stride = (u64)gridDim.x * 256;
    unsigned long long* dst = c->p.top.acc;
    const unsigned long long* src = c->lv[0].acc;
    for (u64 k = (u64)blockIdx.x * 256 + threadIdx.x; k < c->p.top.n; k += stride) dst[k] += src[k];
}

template <int B>
static void plan_steps(const SliderTables* tabs, const ZTables* zt, TTPlanCtl* ctl, StepTable* tab, uint4* tt,
                       u64 mask, unsigned long long* ttstats, int r, int write_grid, int leaf_grid, int up_grid,
                       hipStream_t s) {
    for (int L = 0; L <= r - 3; L++) {
        hipLaunchKernelGGL(k_plan_write, dim3(1), dim3(256), 0, s, ctl, tab, L, tt, mask, ttstats);
        hipLaunchKernelGGL(k_plan_clear, dim3(up_grid), dim3(256), 0, s, ctl, L);
        // the write launch; at the last write level, when its planner chose to
        // fuse the leaf pass (StepTable::fuse_leaf), it leaves at once and the
        // fused launch runs the level instead
        hipLaunchKernelGGL((k_step<B, kTTWaves, kWrite>), dim3(write_grid), dim3(kTWG), 0, s, tabs, zt, tab);
        if (L == r - 3)
            hipLaunchKernelGGL((k_step_write_leaf<B, kTTMixedWaves>), dim3(leaf_grid), dim3(kTWG), 0, s, tabs, zt,
                               tab);
    }
    // records the leaf level; the leaf launch is empty when the write fused it
    hipLaunchKernelGGL(k_plan_leaf, dim3(1), dim3(64), 0, s, ctl, tab, r - 2, ttstats);
    hipLaunchKernelGGL((k_step<B, kTTLeafWaves, kLeaf>), dim3(leaf_grid), dim3(kTWG), 0, s, tabs, zt, tab);
    for (int L = r - 3; L >= 0; L--) {
        hipLaunchKernelGGL(k_plan_up, dim3(1), dim3(64), 0, s, ctl, tab, L, ttstats);
        hipLaunchKernelGGL(k_up_step, dim3(up_grid), dim3(256), 0, s, tab, tt, mask, 0);
    }
    hipLaunchKernelGGL(k_plan_commit, dim3(up_grid), dim3(256), 0, s, ctl);
}

hipError_t capture_plan_pass(int backend, const SliderTables* tabs, const ZTables* zt, TTPlanCtl* ctl,
                             StepTable* tab, const TTPlanParams* h_params, TTPlanStatus* h_status, uint4* tt,
                             u64 mask, unsigned long long* ttstats, int r, int write_grid, int leaf_grid,
                             int up_grid, hipStream_t s) {
    if (r < 3 || r > kPlanMaxLevels - 1) return hipErrorInvalidValue;
    hipError_t e = hipMemcpyAsync(&ctl->p, h_params, sizeof(TTPlanParams), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    switch (backend) {
    case kKS: plan_steps<kKS>(tabs, zt, ctl, tab, tt, mask, ttstats, r, write_grid, leaf_grid, up_grid, s); break;
    case kPlain: plan_steps<kPlain>(tabs, zt, ctl, tab, tt, mask, ttstats, r, write_grid, leaf_grid, up_grid, s); break;
    case kFancy: plan_steps<kFancy>(tabs, zt, ctl, tab, tt, mask, ttstats, r, write_grid, leaf_grid, up_grid, s); break;
    default: plan_steps<kByte>(tabs, zt, ctl, tab, tt, mask, ttstats, r, write_grid, leaf_grid, up_grid, s); break;
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(h_status, &ctl->st, sizeof(TTPlanStatus), hipMemcpyDeviceToHost, s);
}

// ---- launchers ---------------------------------------------------------------
template <int M>
static void* pick_mode(int backend) {
    switch (backend) {
    case kKS: return reinterpret_cast<void*>(&k_expand_tt<kKS, M>);
    case kPlain: return reinterpret_cast<void*>(&k_expand_tt<kPlain, M>);
    case kFancy: return reinterpret_cast<void*>(&k_expand_tt<kFancy, M>);
    default: return reinterpret_cast<void*>(&k_expand_tt<kByte, M>);
    }
}
static void* pick(int backend, int mode) { return mode == kLeaf ? pick_mode<kLeaf>(backend) : pick_mode<kWrite>(backend); }

static int tt_bpc(void* f) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kTWG, 0) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();  // a failed query must not surface as a later call's error
        per_cu = 1;
    }
    return per_cu;
}
int expand_tt_max_bpc(int backend) { return tt_bpc(pick(backend, kWrite)); }
int expand_tt_leaf_bpc(int backend) { return tt_bpc(pick(backend, kLeaf)); }

hipError_t launch_expand_tt(int backend, const SliderTables* tabs, const ZTables* zt, const TTArgs& a, int grid,
                            hipStream_t s) {
    if (!a.n_in) return hipSuccess;
    const u32 tiles = (a.n_in + kTTP - 1) / kTTP;
    if ((u32)grid > tiles) grid = (int)tiles;
    TTArgs args = a;
    void* kargs[] = {(void*)&tabs, (void*)&zt, (void*)&args};
    return hipLaunchKernel(pick(backend, a.mode), dim3(grid), dim3(kTWG), kargs, 0, s);
}

hipError_t launch_roots_tt(int backend, const SliderTables* tabs, Board* roots, const u64* k1, const u64* k2, u32 n,
                           int depth, const uint4* tt, u64 mask, unsigned long long* acc, unsigned long long* total,
                           unsigned long long* stats, hipStream_t s) {
    if (!n) return hipSuccess;
    const dim3 g((n + 255) / 256), b(256);
    switch (backend) {
    case kKS:
        hipLaunchKernelGGL(k_roots_tt<kKS>, g, b, 0, s, tabs, roots, k1, k2, n, depth, tt, mask, acc, total, stats);
        break;
    case kPlain:
        hipLaunchKernelGGL(k_roots_tt<kPlain>, g, b, 0, s, tabs, roots, k1, k2, n, depth, tt, mask, acc, total, stats);
        break;
    case kFancy:
        hipLaunchKernelGGL(k_roots_tt<kFancy>, g, b, 0, s, tabs, roots, k1, k2, n, depth, tt, mask, acc, total, stats);
        break;
    default:
        hipLaunchKernelGGL(k_roots_tt<kByte>, g, b, 0, s, tabs, roots, k1, k2, n, depth, tt, mask, acc, total, stats);
        break;
    }
    return hipGetLastError();
}


hipError_t launch_upsweep(const UpSeg& u, uint4* tt, u64 mask, unsigned long long* stores, hipStream_t s) {
    const u32 items = u.n + u.ndup;
    if (!items) return hipSuccess;
    hipLaunchKernelGGL(k_upsweep, dim3((items + 255) / 256), dim3(256), 0, s, u, tt, mask, stores);
    return hipGetLastError();
}

__device__ __forceinline__ int entry_depth(const uint4& e) { return (int)(e.w >> 24); }  // w1 >> 56

// Folds entry e into a bucket held in registers, keeping its two deepest
// entries (empty entries have depth 0).
__device__ __forceinline__ void fold_entry(uint4& t0, uint4& t1, const uint4& e) {
    if (!(e.x | e.y | e.z | e.w)) return;
    const int d = entry_depth(e);
    if (d >= entry_depth(t0)) {
        t1 = t0;
        t0 = e;
    } else if (d >= entry_depth(t1)) {
        t1 = e;
    }
}

// Saved buckets [first, first + nb) of a table of src_buckets buckets -> this
// table of mask + 1 buckets (both powers of two; this table is cleared
// before the first chunk).  A bucket's index is the low bits of its
// entries' first key:
//  * same size or larger: source bucket s may hold keys of any target bucket
//    s + k * src_buckets, so it is copied to each of them (the entries are
//    still verified by their second key);
//  * smaller: source buckets s, s + N, s + 2N, ... (N = mask + 1) fold into
//    target s & mask.  One thread owns one target and folds all of the
//    chunk's sources of it, so no two threads write one bucket.
// The host keeps first + nb <= src_buckets.
__global__ __launch_bounds__(256) void k_tt_refit(const uint4* src, u64 first, u64 nb, u64 src_buckets, uint4* tt,
                                                  u64 mask) {
    const u64 n = mask + 1, stride = (u64)gridDim.x * 256;
    u64 i = (u64)blockIdx.x * 256 + threadIdx.x;
    if (n >= src_buckets) {
        const u64 total = nb * (n / src_buckets);
        for (; i < total; i += stride) {
            const u64 j = i % nb, k = i / nb;
            uint4* t = tt + 2 * (first + j + k * src_buckets);
            t[0] = src[2 * j];
            t[1] = src[2 * j + 1];
        }
        return;
    }
    const u64 owners = nb < n ? nb : n;
    for (; i < owners; i += stride) {
        uint4* t = tt + 2 * ((first + i) & mask);
        uint4 t0 = t[0], t1 = t[1];
        for (u64 j = i; j < nb; j += n) {
            fold_entry(t0, t1, src[2 * j]);
            fold_entry(t0, t1, src[2 * j + 1]);
        }
        t[0] = t0;
        t[1] = t1;
    }
}

hipError_t launch_tt_refit(const uint4* src, u64 first, u64 nb, u64 src_buckets, uint4* tt, u64 mask, int n_cu,
                           hipStream_t s) {
    if (!nb) return hipSuccess;
    hipLaunchKernelGGL(k_tt_refit, dim3(n_cu * 8), dim3(256), 0, s, src, first, nb, src_buckets, tt, mask);
    return hipGetLastError();
}

}  // namespace pf
