// kernels.h -- host-callable launchers for the HIP kernels in kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "convert.h"
#include "movegen.h"

namespace pf {

// Device view of the slider tables (pointers into one device blob).
struct DevTables {
    const SliderTables* t;
};

// One breadth-first level: parents `in[0..n_in)` with their root ids.
struct LevelArgs {
    const Board* in;
    const u32* in_root;
    u32 n_in;
    Board* out;          // children (expand-and-write mode only)
    u32* out_root;
    u64 out_cap;         // capacity of out / out_root in nodes
    u32* out_count;      // device bump counter for out
    unsigned long long* root_sums;  // per-root accumulators (nullptr: skip)
    unsigned long long* total;      // sum over all children (nullptr: skip)
    u32* tile_ctr;       // persistent tile queue counter (zero before launch)
    u32* err;            // set non-zero on capacity overflow
    u32 tile_parents;    // parents per tile, at most the shape's (0: the shape's); a small level
                         // uses fewer, so that it spreads over more workgroups
};

// Parents per tile for a level of n parents on `grid` workgroups of a shape
// with at most tp_max parents per tile: enough tiles for every workgroup.
PF_HD u32 level_tile_parents(u64 n, u64 grid, u32 tp_max) {
    const u64 g = grid ? grid : 1, t = (n + g - 1) / g;
    return (u32)(t < 1 ? 1 : t > tp_max ? tp_max : t);
}

enum ExpandMode : int { kLeaf = 0, kWrite = 1 };

// Occupancy-derived grid size for the expand kernel (workgroups); small = 1
// for the shape the level loop uses on small levels.
int expand_grid(int backend, int mode, int small, int n_cu);

hipError_t launch_init_roots(int backend, const DevTables& dt, Board* roots, u32 n,
                             unsigned long long* root_sums, hipStream_t s);
hipError_t launch_expand(int backend, int mode, int small, const DevTables& dt, const LevelArgs& a,
                         int grid, hipStream_t s);
hipError_t launch_range_count(const Board* b, u32 n, unsigned long long* out, hipStream_t s);
// Roots already in device memory (reference layout) -> boards + root ids.
hipError_t launch_pos_to_board(const pf_pos48* in, u32 n, Board* out, u32* rid, hipStream_t s);

// Whole-batch breadth-first perft in one cooperative launch (grid barrier
// between levels).  Levels 1..depth-2 are bump-allocated in `arena`.
struct PersistCtl {
    u32 tile_ctr[20];
    u32 out_count[20];
    u32 err;
    u32 pad[3];
    unsigned long long totals[20];  // totals[L] = number of nodes of level L
    unsigned long long lvl_off[20];  // stream-ordered form: arena offset of level L
};
struct PersistArgs {
    Board* roots;
    u32* root_ids;
    u32 n_roots;
    int depth;
    char* arena;
    u64 arena_bytes;
    PersistCtl* ctl;
    unsigned long long* sums;
};
// Tile-shape variants of the persistent kernel (kernels.hip, PCfg) and the
// most workgroups per CU that can be resident for one of them.
constexpr int kPersistVariants = 3;
int persistent_max_bpc(int backend, int variant);
hipError_t launch_persistent(int backend, int variant, const DevTables& dt, const PersistArgs& a, int grid,
                             hipStream_t s);
// The same batch as one launch per level on stream s (levels sized and placed
// on the device; no host round trip): the root count, then depth - 1 level
// launches of grid[kWrite] / grid[kLeaf] workgroups of the level kernel's
// form variant[kWrite] / variant[kLeaf] (stream.hip: tile shape * 3 + waves
// per SIMD target; the byte backend has one form).
constexpr int kStreamVariants = 6;
constexpr int kStreamWriteVariants = 5;  // the write kernel has no five-wave 256 x 4096 form
constexpr int kStreamByteLds = 6;        // byte backend only: its tables staged in LDS (both modes)
constexpr int kStreamLeafDefault = 5;   // 256 x 4096 tile, five waves per SIMD
constexpr int kStreamWriteDefault = 3;  // 256 x 4096 tile, no waves target
int stream_grid(int backend, int mode, int variant, int n_cu);
hipError_t launch_stream(int backend, const DevTables& dt, const PersistArgs& a, const int grid[2],
                         const int variant[2], hipStream_t s);

}  // namespace pf
