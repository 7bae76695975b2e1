// engine.h -- host runtime of the perft engine: device context, frontier
// arena, level driver with depth-first chunking.
#pragma once
#include <mutex>
#include <vector>

#include "../../include/perft_gpu_amd.h"
#include "convert.h"
#include "kernels.h"
#include "tt_kernels.h"

namespace pf {

constexpr u64 kAlignTT = 256;
enum TuneWhat : int { kTuneSinglePass = 1, kTuneTable = 2 };

struct Engine {
    std::recursive_mutex mu;
    int device = 0;
    int backend = kKS;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    SliderTables* d_tab = nullptr;

    // frontier arena (stack allocator)
    char* arena = nullptr;
    u64 arena_bytes = 0, arena_top = 0;

    // control block: [0] err, [1] tile counter, [2] out counter
    u32* d_ctrl = nullptr;
    unsigned long long* d_total = nullptr;   // range_count output
    unsigned long long* d_sums = nullptr;    // final per-root sums
    unsigned long long* d_wsums = nullptr;   // per-root sums of expand-and-write launches
    u64 sums_cap = 0;
    unsigned long long* h_pinned = nullptr;  // small pinned readback buffer

    int grid[4][2] = {};        // level-loop launches on large levels (wide tile, five waves per SIMD)
    int grid_small[4][2] = {};  // the same on small levels (256 x 8192 tile)
    int pgrid[4] = {};
    int sgrid[4][2] = {};  // stream-ordered level launches: [backend][kLeaf / kWrite]
    // launch parameters per backend, chosen by autotune() or a tune file
    // (0 blocks per CU = as many as can be resident)
    struct Tune {
        int p_variant = 0;    // persistent kernel tile shape (kernels.h)
        int p_bpc = 0;        // persistent kernel workgroups per CU
        int tt_bpc = 0;       // TT write-level kernel workgroups per CU
        int single_pass = 0;  // table-free batches under PF_MODE_AUTO: 0 stream-ordered levels, 1 cooperative
        int s_leaf = kStreamLeafDefault;    // stream-ordered leaf-level kernel form (stream.hip)
        int s_write = kStreamWriteDefault;  // stream-ordered write-level kernel form
        int tt_leaf_bpc = 0;  // TT leaf-level kernel workgroups per CU
        int tt_chunk_pct = 50;  // TT level driver: a chunk's next level may take this share of the free arena
    };
    Tune tune[4];
    PersistCtl* d_pctl = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    pf_stats st = {};
    double run_expand_ms = 0;
    u64 run_expand_launches = 0;
    bool timing = true;

    int depth = 0;  // depth of the current run
    bool serving = false;  // the serving loop owns the arena

    // transposition table (tt.h): 2^k buckets of two 16-byte entries
    uint4* d_tt = nullptr;
    u64 tt_buckets = 0;
    ZTables h_z;
    ZTables* d_z = nullptr;
    unsigned long long* d_ttstats = nullptr;  // probes, hits, stores, -, duplicates' moves, duplicates
    int tt_grid[4] = {};       // TT write levels (tuned workgroups per CU x CUs)
    int tt_leaf_grid[4] = {};  // TT leaf levels (the leaf kernel's full occupancy)
    bool dedup = true;  // in-level duplicate detection on TT runs
    bool exact_ep = true;  // children flag en passant only when capturable (else after every double push)
    int tt_split = 0;      // log2 of the table's per-depth regions (0: one table for every depth)
    u64 tt_mask_word() const { return (tt_buckets - 1) | ((u64)tt_split << 56); }
    // Device-planned table passes (engine_tt.cpp): a chunk whose subtree is
    // predicted to fit the free arena runs as one replay of a captured graph.
    bool tt_graphs = true;
    TTPlanCtl* d_plan = nullptr;
    StepTable* d_ptab = nullptr;
    TTPlanParams* h_plan = nullptr;   // pinned
    TTPlanStatus* h_pstat = nullptr;  // pinned
    hipGraph_t plan_graph[4][kPlanMaxLevels] = {};
    hipGraphExec_t plan_exec[4][kPlanMaxLevels] = {};
    double plan_bpc[kPlanMaxLevels] = {};  // predicted pass bytes per child of the chunk, by remaining depth

    struct TLevel {  // one breadth-first level with TT bookkeeping (80 B/node)
        Board* b;
        u32* root;
        u64* k1;
        u64* k2;
        u32* parent;
        unsigned long long* acc;
        u32* bsum;  // per 1024-node block: sum of the nodes' move counts (serving levels only)
        // duplicates of the level's nodes: claim table and list (null: none)
        DupEntry* dup;
        u64 dup_n;
        u32* dpar;
        u32* dslot;
        u64 ndup;
    };

    pf_status init(const pf_config& cfg);
    void destroy();
    pf_status ensure_sums(u64 n);
    pf_status perft_levels(const pf_pos48* pos, int n, int depth, u64* out);
    // dev_in / dev_out: `pos` / `out` are device memory of this engine's device
    pf_status perft_persistent(const pf_pos48* pos, int n, int depth, u64* out, bool dev_in = false,
                               bool dev_out = false);
    // the same batch as one launch per level on the stream, sized on the device
    pf_status perft_stream(const pf_pos48* pos, int n, int depth, u64* out, bool dev_in = false,
                           bool dev_out = false);
    // the stream-ordered batch on device buffers with no allocation: its
    // control block, roots and levels in a fixed device array (noalloc.hip)
    pf_status perft_noalloc(const pf_pos48* dpos, int n, int depth, u64* dout);
    pf_status init_tt(u64 bytes, u64 seed);
    void apply_tuning();
    pf_status load_tuning(const char* path);
    pf_status save_tuning(const char* path) const;
    // what: kTuneSinglePass (the table-free single passes) | kTuneTable (the
    // table's level driver, when one is configured)
    pf_status autotune(const pf_pos48* pos, int n, int depth, int what);
    pf_status probe(int kind, u64 bytes, u64 iters, double* out);

  private:
    char* alloc(u64 bytes);
    void release(char* p);
    pf_status expand(int mode, const Board* in, const u32* in_root, u64 n_in, Board* out, u32* out_root,
                     u64 out_cap, unsigned long long* sums, u64* next_total);
    pf_status range_count(const Board* b, u64 n, u64* total);
    pf_status upload_roots(const pf_pos48* pos, int n, Board** roots, u32** rid, bool dev_in = false);
    pf_status descend(const Board* lvl, const u32* rts, u64 n, u64 n_next, int L);
    bool alloc_tlevel(u64 n, TLevel* lv);
    pf_status expand_tt(int mode, const TLevel& lv, u64 off, u64 m, TLevel* nx, u64 cap, int L, u64* written,
                        u64* next_total);
    pf_status descend_tt(const TLevel& lv, u64 off, u64 m, u64 n_next, int L);
    pf_status descend_tt_host(const TLevel& lv, u64 off, u64 m, u64 n_next, int L);
    pf_status run_plan(const TLevel& lv, u64 off, u64 m, u64 n_next, int L);
    pf_status plan_exec_for(int r, hipGraphExec_t* out);
    double plan_bytes_per_child(int r, double bf) const;
  public:
    void drop_plan_graphs();
  private:
  public:
    pf_status perft_levels_tt(const pf_pos48* pos, int n, int depth, u64* out);
    pf_status tt_clear();
    pf_status tt_save(const char* path);
    pf_status tt_load(const char* path);
};

}  // namespace pf
