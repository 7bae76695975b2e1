// tt_kernels.h -- launchers of the transposition-table perft kernels (tt.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "backends_dev.h"
#include "kernels.h"
#include "tt.h"

namespace pf {

// Table probes and stores only for nodes with at least this remaining depth:
// a node of remaining depth 2 costs one count pass over its children, about
// what a random 32-byte probe plus its share of the store costs.
constexpr int kMinTTDepth = 3;

// Parents per tile of the table kernels (a tile's move list is one LDS
// window of 4096 entries); the serving step's tile prefix counts with it too.
constexpr int kTTTileParents = 112;

// In-level duplicate detection, at the moment a child is made (open
// addressing, linear probing).  `key` is the child's second Zobrist key
// (0 = empty slot), claimed by compare-and-swap; the first child of a
// position to claim its slot is written to the level and publishes its
// index in `val`; a later child with the same key (its slot comes from the
// first key, so a match verifies both) is not written -- it joins the
// level's duplicate list as (parent, slot) and takes the claimant's count in
// the upsweep (the reference's writeIndexInHashForDuplicates /
// checkandMarkDuplicates, perft_bb.h:788-846, and perft_bb.h:871-878).
struct alignas(16) DupEntry {
    unsigned long long key;
    unsigned long long val;  // index of the claimant in its level
};
constexpr int kDupProbes = 32;
constexpr int kDupMinPly = 3;  // no two move sequences shorter than 3 plies reach the same position
constexpr u32 kBsumShift = 10;  // block sums of move counts over 1024 consecutive nodes
constexpr u32 kBsumBlock = 1u << kBsumShift;

// Slots of the duplicate table of a level of n nodes: linear probing at a load
// of at most 2/3, indices below 2^32.
PF_HD u64 dup_table_slots(u64 n) {
    const u64 s = n + n / 2 + 64;
    return s > 0xFFFFFFFFull ? 0xFFFFFFFFull : s;
}

struct TTArgs {
    // parents: a contiguous range of one level
    const Board* in;
    const u32* in_root;   // root index | side-to-move colour << 31
    const u64* in_k1;
    const u64* in_k2;
    u32 n_in;
    u32 in_base;          // index of in[0] within its level (parent indices)
    unsigned long long* in_acc;  // parents' accumulators (+= hits or perft(.,2))
    // children (write mode)
    Board* out;
    u32* out_root;
    u64* out_k1;
    u64* out_k2;
    u32* out_parent;
    unsigned long long* out_acc;
    u64 out_cap;
    u32* out_count;
    unsigned long long* total;   // sum of the written children's move counts
    uint4* tt;                   // 2 x 16-byte entries per bucket
    u64 tt_mask;                 // buckets - 1
    int child_depth;             // remaining depth of the children
    int mode;                    // kWrite or kLeaf
    int ep_always;               // children flag en passant after every double push (keys differ)
    u32* tile_ctr;
    u32* err;
    unsigned long long* stats;   // [0] probes, [1] hits
    DupEntry* dup;               // children claim a slot here (null: no duplicate detection)
    u64 dup_n;                   // slots (< 2^32)
    u32* dup_count;              // duplicates listed so far
    u32* dup_par;                // duplicate list: parent index (in its level) ...
    u32* dup_slot;               // ... and the slot its claimant holds; capacity out_cap
    u32* out_bsum;               // per kBsumBlock children: sum of their move counts (null: not kept)
    u32 bsum_n;                  // blocks (cleared by the serving step's start kernel)
};

// One serving step (serve.cpp): expand/leaf segments of many jobs, their
// upsweep segments and newly admitted roots, described in device memory so
// that the step can be replayed from a captured graph.
constexpr int kMaxSegs = 512;
constexpr int kMaxUps = 512;
constexpr int kMaxRoots = 256;
// Serving: leaf passes per step over levels written in the same step.
constexpr int kMaxLate = 16;
// Serving: a step whose levels are small splits them into tiles of as few as
// this many parents (about 30 children each: one child per thread of a
// workgroup), the fewest that still give every tile a workgroup of the launch.
constexpr u32 kFineMinParents = 8;

struct SegOut {
    u32 out_count;
    u32 err;
    unsigned long long total;
    u32 ndup;  // children found to duplicate a written sibling of their level
    u32 pad;
};
struct UpSeg {
    const u32* parent;
    const u64* k1;
    const u64* k2;
    const unsigned long long* acc;
    unsigned long long* parent_acc;
    const DupEntry* dup;         // the level's duplicate list and table
    const u32* dup_par;
    const u32* dup_slot;
    u32 n, ndup;                 // nodes, then duplicates: n + ndup items
    int depth;
    // serving, upsweeps that follow a write in the same step (a late leaf
    // pass's chain): that write's outcome.  No items if it overflowed; with
    // dev_counts, the level is that write's and its n and ndup are there.
    const SegOut* dev;
    u32 dev_counts;
};
struct RootSeg {  // payload travels in the table; the kernel scatters it
    Board b;
    u64 k1, k2;
    u32 word;  // slot | side-to-move colour << 31
    int depth;
    Board* dst_b;
    u64* dst_k1;
    u64* dst_k2;
    u32* dst_w;
    unsigned long long* acc;
};
struct RootOut {
    u32 hit;
    u32 count;
};
// Upsweeps of one step run in kUpRounds dependent launches: round r+1 adds
// levels whose own children were added in round r (a finished leaf level and
// the chain of levels above it complete in one step).
constexpr int kUpRounds = 4;

struct StepTable {
    u32 nseg, nup, nroot, tile_ctr;
    // segments [0, nseg_write) are write levels, [nseg_write, nseg) leaf
    // passes, each kind with its own tile counter (tile_ctr, tile_ctr2)
    u32 nseg_write, tile_ctr2;
    // the serving step's mixed launch: workgroups [0, mix_write_wgs) start on
    // the write tiles, the others on the leaf tiles
    u32 mix_write_wgs;
    // parents per tile of the mixed launch's finely tiled form (its tile
    // prefix counts with it); the other form tiles by kTTTileParents
    u32 mix_tp;
    // planned pass: the last write level runs its leaf pass fused (set by its
    // planner when the level has enough tiles for every workgroup)
    u32 fuse_leaf;
    // serving: leaf passes over levels written in this same step, run after
    // the write launch ("late"): late_pos[k] is the write segment's position in
    // seg[]; their tiles come from tile_ctr3
    u32 nlate, tile_ctr3;
    u32 late_pos[kMaxLate];
    u32 seq;  // serving: the step's number, which the post kernel writes to done last
    unsigned long long* stores;    // table stores of the upsweeps are counted here (null: not counted)
    u32 tile_prefix[kMaxSegs + 1];
    u32 up_prefix[kMaxUps + 1];    // item prefix over all upsweep segments (the upsweep kernel recomputes
                                   // it per round, as some counts are only known on the device)
    u32 up_round[kUpRounds + 1];   // segments [up_round[r], up_round[r+1]) belong to round r
    TTArgs seg[kMaxSegs];
    UpSeg up[kMaxUps];
    RootSeg root[kMaxRoots];
    // device-written part, from out[] on: serving's fetch kernel copies in
    // the used entries before it, its post kernel the used entries from it
    SegOut out[kMaxSegs];          // zeroed by the start kernel for [0, nseg)
    RootOut rout[kMaxRoots];
    unsigned long long acc[kMaxRoots];  // serving: each job slot's root accumulator
    unsigned long long stats[2];        // serving: table probes and hits of the write segments
    u32 done;                           // serving: seq of the last step whose outcomes are all in the host table
    u32 rb_done;                        // serving: number of the last read-back that has reached the host
};

// ---- device-planned table pass ----------------------------------------------
// A chunk of parents whose whole subtree fits the arena runs as one replay of
// a captured graph (engine_tt.cpp): every write level is sized and placed on
// the device by a one-block planner from the counts the previous launch left
// there, so no level returns to the host (the reference's per-level launch
// and device sync, perft_bb.h:1071-1135).  Per write level: plan, clear the
// level's duplicate table, expand; the last write level runs fused with the
// leaf pass over its children; then one upsweep per level on the way back up;
// then the commit.  The pass accumulates the
// parents' counts in an array of its own and adds them to the caller's only
// at the commit, and only if no level overflowed: an overflowed pass leaves
// the caller's accumulators, the arena below it and the table unchanged.
constexpr int kPlanMaxLevels = 18;
// The last write level fuses its leaf pass when it has at least this many
// tiles of parents per workgroup of the launch.
constexpr int kFuseTilesPerWG = 16;
struct TTLevelRec {
    Board* b;
    u32* root;
    u64* k1;
    u64* k2;
    u32* parent;
    unsigned long long* acc;
    DupEntry* dup;
    u64 dup_n;
    u32* dpar;
    u32* dslot;
    u64 n;     // nodes in the level
    u64 ndup;  // duplicates listed with it
    u64 next;  // sum of its nodes' move counts: the size of the level below
};
struct TTPlanParams {  // the host's part, copied in by the graph's first node
    TTLevelRec top;    // the chunk's parents (acc: the caller's accumulators)
    char* arena;       // the pass's memory [arena, arena + arena_bytes)
    u64 arena_bytes;
    int depth;         // remaining depth of the parents (>= 3)
    int ply;           // ply of the parents (duplicate tables from kDupMinPly)
    int dedup;
    int ep_always;     // TTArgs::ep_always of the pass's write levels
    u32 fuse_min_tiles;  // the last write level fuses its leaf pass from this many tiles of parents
};
struct TTPlanStatus {  // copied back by the graph's last node
    u64 peak;          // arena bytes the pass used
    u32 err;           // 2: a level did not fit (the pass had no effect)
    u32 pad;
    unsigned long long made, frontier, dups;
};
struct TTPlanCtl {
    TTPlanParams p;
    TTPlanStatus st;
    TTLevelRec lv[kPlanMaxLevels];
};
// Appends one pass for parents of remaining depth r (3..kPlanMaxLevels) to
// the stream being captured: params H2D from *h_params, the kernels, status
// D2H to *h_status.
hipError_t capture_plan_pass(int backend, const SliderTables* tabs, const ZTables* zt, TTPlanCtl* ctl,
                             StepTable* tab, const TTPlanParams* h_params, TTPlanStatus* h_status, uint4* tt,
                             u64 mask, unsigned long long* ttstats, int r, int write_grid, int leaf_grid,
                             int up_grid, hipStream_t s);

// Serving: the used part of the step table from coherent host memory `h`
// into `d`; the step's outcomes from `d` back into `h`, then h->done = d->seq
// once every outcome is visible to the host.
hipError_t launch_fetch_step(StepTable* d, const StepTable* h, hipStream_t s);
hipError_t launch_post_step(const StepTable* d, StepTable* h, u32 max_jobs, hipStream_t s);
// *flag = v (host memory, system scope) once the stream's earlier work is done.
hipError_t launch_post_flag(u32* flag, u32 v, hipStream_t s);
int step_grid(int backend, int n_cu);       // workgroups of the serving step's mixed launch
int step_max_bpc(int backend, int leaf);  // the planned pass's write (0) / leaf (1) step kernel
// fine: the mixed launch tiles by StepTable::mix_tp parents, not kTTTileParents;
// late: the late leaf launch is recorded (steps with StepTable::nlate > 0)
hipError_t launch_step(int backend, const SliderTables* tabs, const ZTables* zt, StepTable* t, uint4* tt, u64 mask,
                       int grid, int up_grid, int up_rounds, int fine, int late, hipStream_t s);
int expand_tt_max_bpc(int backend);   // most resident workgroups per CU of the TT write-level kernel
int expand_tt_leaf_bpc(int backend);  // the same for the TT leaf-level kernel
hipError_t launch_expand_tt(int backend, const SliderTables* tabs, const ZTables* zt, const TTArgs& a, int grid,
                            hipStream_t s);
// Roots: probe (counted in stats[0], hits in stats[1]) or count their moves.
hipError_t launch_roots_tt(int backend, const SliderTables* tabs, Board* roots, const u64* k1, const u64* k2, u32 n,
                           int depth, const uint4* tt, u64 mask, unsigned long long* acc, unsigned long long* total,
                           unsigned long long* stats, hipStream_t s);
// A finished level: its n nodes and ndup duplicates add their counts to their
// parents (parent_acc null: the roots); nodes are stored in the table.
hipError_t launch_upsweep(const UpSeg& u, uint4* tt, u64 mask, unsigned long long* stores, hipStream_t s);
// A chunk of saved buckets [first, first + nb) of a src_buckets-bucket table
// into this (cleared) table of mask + 1 buckets.
hipError_t launch_tt_refit(const uint4* src, u64 first, u64 nb, u64 src_buckets, uint4* tt, u64 mask, int n_cu,
                           hipStream_t s);

}  // namespace pf
