// serve.cpp -- continuous-batching perft server (no reference counterpart: the
// reference answers one perft per process run, perft.cu:133-137).
//
// Admission is by predicted frontier memory: a request asks for the bytes its
// breadth-first levels are predicted to need (root move count as branching
// factor, the levels down to depth-2 plus a duplicate table, clamped to
// [1 MiB, arena/4]) and is admitted when a job slot is free and a region of
// that size is free in the arena (first fit; a younger request may pass an
// older one that does not fit, unless the older one has waited 50 ms).  A
// job never waits on another job's memory once admitted: where the
// prediction falls short it chunks depth-first inside its own region.  With
// a fixed job_bytes every job gets a region of exactly that size.  A job is
// the chunked level driver of engine_tt.cpp rewritten as an explicit frame
// stack, so that it can advance one GPU action at a time.
//
// One scheduler step = one replay of a captured graph:
//   H2D of the step table's host part -> start kernel (zeroes the segment
//   outcomes; new requests: table probe at full depth, move count; empties
//   the duplicate tables and block sums of this step's write segments) ->
//   step kernel (the write and leaf levels of every active job in one
//   persistent launch over two tile queues; duplicates are listed as they are
//   made) -> up to kUpRounds upsweep kernels (finished levels of every job,
//   each round the levels just above the previous round's; one graph per
//   round count) -> D2H of the table's device part: per-segment outcomes,
//   slot accumulators, table counters.
// Requests that arrive while a step runs join the next step; their work
// overlaps on the device with everyone else's.  All jobs share the lockless
// transposition table.
#include "serve.h"

#include <string.h>

#include "movegen.h"
#include "sliders.h"

#include <algorithm>
#include <chrono>

namespace pf {

static const auto kEpoch = std::chrono::steady_clock::now();
double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - kEpoch).count(); }

static constexpr u64 kTNode = sizeof(Board) + 4 + 8 + 8 + 4 + 8;
static constexpr u64 kAl = 256;
// frame states: START plans a write or leaf level; WRITE / LEAF / UP are in
// flight; WAITCHILD parks a parent under its child frame; CHILD (child done)
// plans the upsweep; UPDONE frees the children and moves on.
enum { ST_START = 0, ST_WRITE, ST_LEAF, ST_WAITCHILD, ST_CHILD, ST_UP, ST_UPDONE, ST_WRITE_LATE };
// A step may hold late leaf passes (the write of a level of remaining depth 2,
// its leaf pass and the upsweeps above it, all in one step) when nothing waits
// for admission and at most this many jobs are active: small steps, whose
// latency it shortens by a step per job.  Busier steps overlap the write and
// the leaf pass of different jobs in the mixed launch instead.
static constexpr size_t kLateMaxJobs = 16;

Server::Server(Engine& e, int max_jobs, u64 job_bytes) : e_(e), max_jobs_(max_jobs), job_bytes_(job_bytes) {}

Server::~Server() { stop(); }

pf_status Server::start() {
    if (running_) return PF_OK;
    if (max_jobs_ < 1 || max_jobs_ > kMaxRoots) return PF_EINVAL;
    job_bytes_ &= ~(kAl - 1);
    if (job_bytes_ && (job_bytes_ * (u64)max_jobs_ > e_.arena_bytes || job_bytes_ < (1u << 20))) return PF_ENOMEM;
    if (e_.arena_bytes < (4u << 20)) return PF_ENOMEM;
    free_mem_.clear();
    free_mem_[0] = e_.arena_bytes & ~(kAl - 1);
    max_region_ = job_bytes_ ? job_bytes_ : ((e_.arena_bytes / 4) & ~(kAl - 1));
    if (hipSetDevice(e_.device) != hipSuccess) return PF_EDEVICE;
    if (hipHostMalloc(&h_tab_, sizeof(StepTable), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess)
        return PF_ENOMEM;
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&h_dev_), h_tab_, 0) != hipSuccess) return PF_EDEVICE;
    if (hipMalloc(&d_tab_, sizeof(StepTable)) != hipSuccess) return PF_ENOMEM;
    // the slot accumulators and the table counters live in the step table's
    // device-written part, which starts at zero
    if (hipMemset(d_tab_, 0, sizeof(StepTable)) != hipSuccess) return PF_EDEVICE;
    slot_acc_ = d_tab_->acc;
    const int n = max_jobs_;
    if (hipMalloc(&root_b_, sizeof(Board) * n) != hipSuccess || hipMalloc(&root_w_, 4 * n) != hipSuccess ||
        hipMalloc(&root_k1_, 8 * n) != hipSuccess || hipMalloc(&root_k2_, 8 * n) != hipSuccess ||
        hipMalloc(&root_p_, 4 * n) != hipSuccess)
        return PF_ENOMEM;
    if (hipMemset(root_p_, 0, 4 * n) != hipSuccess) return PF_EDEVICE;
    memset(h_tab_, 0, sizeof(StepTable));
    jobs_.assign((size_t)n, Job{});
    seg_tmp_.assign((size_t)kMaxSegs, TTArgs{});
    pos_of_.assign((size_t)kMaxSegs, 0u);
    free_slots_.clear();
    for (int i = n - 1; i >= 0; i--) free_slots_.push_back(i);
    grid_ = step_grid(e_.backend, e_.n_cu);
    up_grid_ = e_.n_cu * 4;
    stop_ = false;
    spin_stop_.store(false);
    queued_.store(queue_.size());
    // the serving thread captures the step graph itself and then replays it,
    // so capture and every replay happen on one thread; start() returns the
    // capture's outcome
    {
        std::unique_lock<std::mutex> g(mu_);
        graph_status_ = 1;
        th_ = std::thread([this] { run(); });
        cv_.wait(g, [&] { return graph_status_ != 1; });
        if (graph_status_ != PF_OK) {
            const pf_status s = (pf_status)graph_status_;
            g.unlock();
            th_.join();
            return s;
        }
        running_ = true;  // under the lock that submit() reads it with
    }
    return PF_OK;
}

// One graph per form of the mixed launch (fine), presence of late leaf passes
// (late) and number of upsweep rounds (0..kUpRounds): a step replays the one
// that fits it, so a small step does not pay for launches it leaves empty.
pf_status Server::build_graph(int fine, int late, int rounds) {
    hipStream_t st = e_.stream;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return PF_EDEVICE;
    // the used part of the host's table in, the outcomes (segments, roots,
    // slot accumulators, table counters) out, each by a kernel that reads or
    // writes the coherent host table: no copy node between the host and the
    // step, and a small step moves only what it uses
    hipError_t e = launch_fetch_step(d_tab_, h_dev_, st);
    if (e == hipSuccess)
        e = launch_step(e_.backend, e_.d_tab, e_.d_z, d_tab_, e_.d_tt, e_.tt_buckets ? e_.tt_mask_word() : 0, grid_,
                        up_grid_, rounds, fine, late, st);
    if (e == hipSuccess) e = launch_post_step(d_tab_, h_dev_, (u32)max_jobs_, st);
    // the capture is ended whatever happened inside it, so the stream leaves capture mode
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(st, &g);
    if (e != hipSuccess || ee != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return PF_EDEVICE;
    }
    graph_[fine][late][rounds] = g;
    if (hipGraphInstantiate(&exec_[fine][late][rounds], g, nullptr, nullptr, 0) != hipSuccess) return PF_EDEVICE;
    return PF_OK;
}

// Waits until the word *w of the host table holds v. The stream's last kernel
// writes it after everything the host needs has reached the host, a few
// microseconds before the stream's completion would wake a waiting thread, so
// the serving thread watches the word (a step's done, a read-back's rb_done).
// Work still running after kSpinWaitS is waited for by hipStreamSynchronize,
// which also reports a failed launch.
bool Server::wait_word(const u32* w, u32 v) {
    constexpr double kSpinWaitS = 0.001;
    const double t0 = now_s();
    while (__atomic_load_n(w, __ATOMIC_ACQUIRE) != v) {
        if (now_s() - t0 > kSpinWaitS)
            return hipStreamSynchronize(e_.stream) == hipSuccess && __atomic_load_n(w, __ATOMIC_ACQUIRE) == v;
        __builtin_ia32_pause();
    }
    return true;
}

// With no job running, the serving thread watches for a submission for up to
// kIdleSpinS before it sleeps on the condition variable, so that a request
// arriving meanwhile starts without the wake-up of a sleeping thread.
void Server::idle_spin() {
    constexpr double kIdleSpinS = 0.020;
    const double t0 = now_s();
    while (queued_.load(std::memory_order_acquire) == 0 && !spin_stop_.load(std::memory_order_acquire) &&
           now_s() - t0 < kIdleSpinS)
        __builtin_ia32_pause();
}

// Workgroups of the mixed step launch that start on the write tiles: a share
// of the workgroups that get a tile (those past the step's tile count leave
// at once) proportional to the write tiles' part of the step's work, a write
// tile costing about as much as kWriteTileCost leaf tiles (C5's 4096
// positions at perft 6 write 1.1e10 children/s and count 5.1e10 in leaf
// passes).  Whichever queue empties first, its workgroups move to the other.
static constexpr double kWriteTileCost = 4.6;
u32 Server::mix_write_wgs(u32 write_tiles, u32 leaf_tiles) const {
    if (!write_tiles) return 0;
    const u32 active = std::min<u32>((u32)grid_, write_tiles + leaf_tiles);
    const double w = kWriteTileCost * write_tiles;
    const u32 n = (u32)((double)active * w / (w + leaf_tiles) + 0.5);
    return std::max(1u, std::min(n, active));
}

void Server::stop() {
    if (running_) {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        spin_stop_.store(true, std::memory_order_release);
        cv_.notify_all();
        if (th_.joinable()) th_.join();
        std::lock_guard<std::mutex> g(mu_);
        running_ = false;
    }
    for (int f = 0; f < 2; f++)
        for (int l = 0; l < 2; l++)
            for (int r = 0; r <= kUpRounds; r++) {
                if (exec_[f][l][r]) hipGraphExecDestroy(exec_[f][l][r]);
                if (graph_[f][l][r]) hipGraphDestroy(graph_[f][l][r]);
                exec_[f][l][r] = nullptr;
                graph_[f][l][r] = nullptr;
            }
    hipFree(d_tab_);
    hipFree(root_b_);
    hipFree(root_w_);
    hipFree(root_k1_);
    hipFree(root_k2_);
    hipFree(root_p_);
    hipHostFree(h_tab_);
    if (h_rb_) hipHostFree(h_rb_);
    h_rb_ = nullptr;
    rb_cap_ = 0;
    d_tab_ = nullptr; root_b_ = nullptr; root_w_ = nullptr; root_k1_ = nullptr; root_k2_ = nullptr;
    root_p_ = nullptr; slot_acc_ = nullptr; h_tab_ = nullptr; h_dev_ = nullptr;
}

// Predicted peak frontier bytes of a request: the written levels are plies
// 1..depth-2 (the last is expanded in leaf mode), each with 80 B per node plus
// 32 B of duplicate table and list and a block-sum share, grown by the root's
// move count per ply (held in [16, 48]) and a quarter of headroom.
u64 Server::predict_bytes(const pf_pos48& pos, int depth) const {
    if (job_bytes_) return job_bytes_;
    const u64 c = count_moves(to_board(pos), KSBackend{});
    const double bf = (double)std::min<u64>(48, std::max<u64>(16, c));
    // (duplicate elimination makes the deep levels smaller, but a region that
    // proves too small costs a chain of chunked steps at every level below,
    // so no credit is taken for it)
    double lvl = (double)c, sum = 0;
    for (int k = 1; k <= depth - 2; k++) {
        sum += lvl;
        lvl *= bf;
    }
    double bytes = sum * (double)(kTNode + 33) * 1.25 + 16.0 * kAl;
    const double lo = (double)(1u << 20), hi = (double)max_region_;
    bytes = std::min(hi, std::max(lo, bytes));
    return ((u64)bytes + kAl - 1) & ~(kAl - 1);
}

bool Server::region_alloc(u64 bytes, u64* off) {
    for (auto it = free_mem_.begin(); it != free_mem_.end(); ++it) {
        if (it->second < bytes) continue;
        *off = it->first;
        const u64 rest = it->second - bytes, at = it->first + bytes;
        free_mem_.erase(it);
        if (rest) free_mem_[at] = rest;
        return true;
    }
    return false;
}

// With nothing waiting for admission (and regions not fixed by job_bytes), a
// job whose level would chunk takes up to `extra` bytes from the free block
// that starts where its region ends, if
// there is one: the last deep requests of a burst then run in a few large
// chunks instead of a long tail of small steps.  The rest of the block stays
// free for later arrivals; what was taken returns with the region when the
// job finishes.
bool Server::grow_region(Job& j, u64 extra) {
    auto it = free_mem_.find(j.base + j.cap);
    if (it == free_mem_.end() || !extra) return false;
    const u64 take_b = std::min(it->second, (extra + kAl - 1) & ~(kAl - 1));
    const u64 rest = it->second - take_b, at = it->first + take_b;
    free_mem_.erase(it);
    if (rest) free_mem_[at] = rest;
    j.cap += take_b;
    return true;
}

void Server::region_free(u64 off, u64 bytes) {
    auto it = free_mem_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_mem_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_mem_.erase(nx);
    }
    if (it != free_mem_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_mem_.erase(it);
        }
    }
}

pf_status Server::submit(const pf_pos48& pos, int depth, u64 tag) {
    if (depth < 0 || depth > 16) return PF_EINVAL;
    Job j;
    j.tag = tag;
    j.depth = depth;
    j.pos = pos;
    j.t_submit = now_s();
    j.want = predict_bytes(pos, depth);
    {
        std::lock_guard<std::mutex> g(mu_);
        if (!running_ || stop_) return PF_EBUSY;
        if (queue_.size() > 1000000) return PF_EBUSY;
        queue_.push_back(std::move(j));
        queued_.store(queue_.size(), std::memory_order_release);
    }
    cv_.notify_all();
    return PF_OK;
}

int Server::poll(pf_result* out, int max) {
    std::lock_guard<std::mutex> g(mu_);
    const int n = std::min<int>(max, (int)done_.size());
    for (int i = 0; i < n; i++) out[i] = done_[(size_t)i];
    done_.erase(done_.begin(), done_.begin() + n);
    return n;
}

void Server::counters(u64* steps, u64* segs, double* busy, u64* dups, u64* probes, u64* hits) {
    std::lock_guard<std::mutex> g(mu_);
    *steps = steps_;
    *segs = segs_;
    *busy = busy_;
    *dups = dups_;
    *probes = probes_ + root_probes_;
    *hits = hits_ + root_hits_;
}

void Server::finish(Job& j, u64 nodes, int status) {
    pf_result r;
    r.tag = j.tag;
    r.nodes = nodes;
    r.status = status;
    r.depth = j.depth;
    r.submit_s = j.t_submit;
    r.done_s = now_s();
    j.done = true;
    j.st.clear();
    {
        std::lock_guard<std::mutex> g(mu_);
        done_.push_back(r);
        if (j.slot >= 0) free_slots_.push_back(j.slot);
    }
    if (j.cap) region_free(j.base, j.cap);
    j.cap = 0;
    j.slot = -1;
}

char* Server::take(Job& j, u64 bytes) {
    bytes = (bytes + kAl - 1) & ~(kAl - 1);
    if (j.top + bytes > j.cap) return nullptr;
    char* p = e_.arena + j.base + j.top;
    j.top += bytes;
    return p;
}

bool Server::alloc_level(Job& j, u64 n, Engine::TLevel* lv) {
    lv->b = reinterpret_cast<Board*>(take(j, n * sizeof(Board)));
    lv->root = reinterpret_cast<u32*>(take(j, n * 4));
    lv->k1 = reinterpret_cast<u64*>(take(j, n * 8));
    lv->k2 = reinterpret_cast<u64*>(take(j, n * 8));
    lv->parent = reinterpret_cast<u32*>(take(j, n * 4));
    lv->acc = reinterpret_cast<unsigned long long*>(take(j, n * 8));
    return lv->b && lv->root && lv->k1 && lv->k2 && lv->parent && lv->acc;
}

// Next chunk of a chunking frame f: parents [f.off, f.off + f.m) whose
// children fit in `budget` nodes; *cap = the children they can have.  From a
// block boundary it takes whole 1024-parent blocks by their exact move sums
// (the level's block sums, copied to the host when the level first chunks);
// inside a block, or when one block alone is too big, as many parents as fit
// at 218 children each, up to the next block boundary.
bool Server::chunk(Job& j, Frame& f, u64 budget, u64* cap) {
    (void)j;
    if (f.lv.bsum && (f.off & (kBsumBlock - 1)) == 0) {
        const u64 nblk = (f.end + kBsumBlock - 1) >> kBsumShift;
        if (f.hbsum.size() != nblk) {
            // through a pinned staging buffer, with a flag after the copy that
            // the serving thread watches (no sleeping wait per read-back)
            if (nblk > rb_cap_) {
                if (h_rb_) hipHostFree(h_rb_);
                h_rb_ = nullptr;
                rb_cap_ = 0;
                const u64 cap = std::max<u64>(nblk, 1u << 16);
                if (hipHostMalloc(reinterpret_cast<void**>(&h_rb_), cap * 4, hipHostMallocDefault) != hipSuccess) {
                    h_rb_ = nullptr;
                    return false;
                }
                rb_cap_ = cap;
            }
            const u32 v = ++rb_seq_;
            if (hipMemcpyAsync(h_rb_, f.lv.bsum, nblk * 4, hipMemcpyDeviceToHost, e_.stream) != hipSuccess ||
                launch_post_flag(&h_dev_->rb_done, v, e_.stream) != hipSuccess || !wait_word(&h_tab_->rb_done, v))
                return false;
            f.hbsum.assign(h_rb_, h_rb_ + nblk);
        }
        u64 m = 0, sum = 0;
        for (u64 b = f.off >> kBsumShift; b < nblk; b++) {
            if (sum + f.hbsum[(size_t)b] > budget) break;
            sum += f.hbsum[(size_t)b];
            m += std::min<u64>(kBsumBlock, f.end - (b << kBsumShift));
        }
        if (m) {
            f.m = m;
            *cap = sum;
            return true;
        }
    }
    u64 lim = f.end - f.off;
    if (f.lv.bsum) lim = std::min(lim, (((f.off >> kBsumShift) + 1) << kBsumShift) - f.off);
    f.m = std::min(std::max<u64>(1, budget / 218), lim);
    *cap = f.m * 218;
    return true;
}

// Frames st[k], st[k-1], ... have complete child levels: schedule their
// upsweeps in rounds 0, 1, ... of this step, up to a frame that still has
// chunks to run (its next chunk reuses the memory being upswept, so it waits
// for the next step).  dev_seg >= 0: st[k]'s child level is written in this
// step by segment dev_seg, whose outcome holds its counts.  Returns the
// number scheduled.
int Server::chain_up(Job& j, int k, int dev_seg) {
    int n = 0;
    for (; k >= 0 && n < kUpRounds; k--) {
        if (nup_planned_ >= (u32)kMaxUps) break;
        Frame& p = j.st[(size_t)k];
        UpSeg u = {};
        u.parent = p.nx.parent;
        u.k1 = p.nx.k1;
        u.k2 = p.nx.k2;
        u.acc = p.nx.acc;
        u.parent_acc = p.lv.acc;
        u.dup = p.nx.dup;
        u.dup_par = p.nx.dpar;
        u.dup_slot = p.nx.dslot;
        u.n = (u32)p.written;
        u.ndup = (u32)p.nx.ndup;
        u.depth = j.depth - p.L - 1;
        if (dev_seg >= 0) {  // st[k]'s child level is written this step: every count of the chain waits on it
            u.dev = &d_tab_->out[dev_seg];
            u.dev_counts = n == 0 ? 1u : 0u;
        }
        up_[(size_t)n].push_back(u);
        nup_planned_++;
        n++;
        if (p.chunking && p.off + p.m < p.end) break;
    }
    return n;
}

// After a step that upswept the child levels of the top n frames (chain_up):
// free those levels, continue a chunked frame or pop finished ones.  Returns
// true when the job's stack emptied (its root count is final).
bool Server::after_chain(Job& j, int n) {
    for (int i = 0; i < n; i++) {
        Frame& p = j.st.back();
        if (p.nx.b) j.top = p.mark;
        p.nx = Engine::TLevel{};
        if (p.chunking && p.off + p.m < p.end) {
            p.off += p.m;
            p.state = ST_START;
            return false;
        }
        j.st.pop_back();
        if (j.st.empty()) return true;
    }
    if (j.st.empty()) return true;
    j.st.back().state = ST_CHILD;  // its child level is complete, not yet upswept
    return false;
}

// Advance job j to its next GPU action (an expand/leaf segment, possibly with
// a chain of upsweeps, or a chain of upsweeps alone).  Returns true if it
// planned one; finishes the job when its frame stack empties.
bool Server::plan(Job& j) {
    StepTable& t = *h_tab_;
    for (;;) {
        if (j.st.empty()) {
            finish(j, h_tab_->acc[j.slot], PF_OK);
            return false;
        }
        Frame& f = j.st.back();
        const int r = j.depth - f.L;
        if (f.state == ST_START) {
            if (f.m == 0) {
                j.st.pop_back();
                if (!j.st.empty()) j.st.back().state = ST_CHILD;
                continue;
            }
            if (t.nseg >= (u32)kMaxSegs) return false;
            TTArgs a = {};
            a.in = f.lv.b + f.off;
            a.in_root = f.lv.root + f.off;
            a.in_k1 = f.lv.k1 + f.off;
            a.in_k2 = f.lv.k2 + f.off;
            a.n_in = (u32)f.m;
            a.in_base = (u32)f.off;
            a.in_acc = f.lv.acc + f.off;
            a.tt = e_.d_tt;
            a.tt_mask = e_.tt_buckets ? e_.tt_mask_word() : 0;
            a.ep_always = e_.exact_ep ? 0 : 1;
            a.child_depth = r - 1;
            a.stats = d_tab_->stats;
            const u32 si = t.nseg;
            a.out_count = &d_tab_->out[si].out_count;
            a.err = &d_tab_->out[si].err;
            a.tile_ctr = nullptr;
            j.chain = 0;
            if (r == 2) {
                // leaf pass over this level; the levels above it that it
                // completes are upswept in the same step
                a.mode = kLeaf;
                f.state = ST_LEAF;
                f.nx = Engine::TLevel{};
                j.chain = chain_up(j, (int)j.st.size() - 2);
            } else {
                // children at ply L+1 get a duplicate table and list (24 + 8 B
                // per child), which live with their level
                const bool dd = e_.dedup && e_.d_tt && f.L + 1 >= kDupMinPly;  // keys exist only with a table
                // a level that the leaf pass will finish (remaining depth 2) is
                // never chunked: it needs no block sums and nothing below it, so
                // it may take all the free memory; others keep 4 B of block sum
                // per 1024 children (charged as 1 B per child) and leave a quarter
                const bool leaf_next = r - 1 == 2;
                const u64 node_b = kTNode + (dd ? 32 : 0) + (leaf_next ? 0 : 1);
                u64 free_b = j.cap - j.top;
                u64 cap;
                u64 room = leaf_next ? free_b : free_b / 2 + free_b / 4;
                const u64 need = f.n_next * node_b + 11 * kAl;  // the whole level, unchunked
                const u64 want_free = leaf_next ? need : need / 3 * 4 + kAl;
                if (!job_bytes_ && queue_idle_ && (f.chunking || need > room) && want_free > free_b &&
                    grow_region(j, want_free - free_b)) {
                    free_b = j.cap - j.top;
                    room = leaf_next ? free_b : free_b / 2 + free_b / 4;
                }
                if (!f.chunking && f.n_next * node_b + 11 * kAl <= room) {
                    cap = f.n_next;
                } else {
                    if (!f.chunking) {
                        f.chunking = true;
                        f.end = f.off + f.m;
                    }
                    const u64 budget = leaf_next ? free_b - std::min<u64>(free_b, 16 * kAl)
                                                 : free_b / 2 - std::min<u64>(free_b / 2, 2 * kAl);
                    if (!chunk(j, f, budget / node_b, &cap)) return false;
                    a.n_in = (u32)f.m;
                }
                f.mark = j.top;
                if (!alloc_level(j, cap, &f.nx)) {
                    j.top = f.mark;
                    return false;  // cannot happen with the half-free rule; wait
                }
                if (!leaf_next) {
                    f.nx.bsum = reinterpret_cast<u32*>(take(j, ((cap + kBsumBlock - 1) >> kBsumShift) * 4));
                    if (!f.nx.bsum) {
                        j.top = f.mark;
                        return false;
                    }
                }
                if (dd && cap >= 2) {
                    const u64 dn = dup_table_slots(cap);
                    const u64 top0 = j.top;
                    f.nx.dup = reinterpret_cast<DupEntry*>(take(j, dn * sizeof(DupEntry)));
                    f.nx.dpar = reinterpret_cast<u32*>(take(j, cap * 4));
                    f.nx.dslot = reinterpret_cast<u32*>(take(j, cap * 4));
                    if (f.nx.dup && f.nx.dpar && f.nx.dslot) {
                        f.nx.dup_n = dn;
                        a.dup = f.nx.dup;
                        a.dup_n = dn;
                        a.dup_count = &d_tab_->out[si].ndup;
                        a.dup_par = f.nx.dpar;
                        a.dup_slot = f.nx.dslot;
                    } else {  // without room the level simply keeps its duplicates
                        j.top = top0;
                        f.nx.dup = nullptr;
                        f.nx.dpar = f.nx.dslot = nullptr;
                    }
                }
                f.cap = cap;
                a.mode = kWrite;
                // children of remaining depth 2 in a small step: their leaf pass
                // and the upsweeps above them run in this step too
                if (r == 3 && late_ok_ && t.nlate < (u32)kMaxLate && nup_planned_ < (u32)kMaxUps) {
                    late_src_[t.nlate++] = si;
                    f.state = ST_WRITE_LATE;
                    j.chain = chain_up(j, (int)j.st.size() - 1, (int)si);
                }
                a.out = f.nx.b;
                a.out_root = f.nx.root;
                a.out_k1 = f.nx.k1;
                a.out_k2 = f.nx.k2;
                a.out_parent = f.nx.parent;
                a.out_acc = f.nx.acc;
                a.out_cap = cap;
                a.out_bsum = f.nx.bsum;
                a.bsum_n = f.nx.bsum ? (u32)((cap + kBsumBlock - 1) >> kBsumShift) : 0u;
                a.total = &d_tab_->out[si].total;
                if (f.state != ST_WRITE_LATE) f.state = ST_WRITE;
            }
            t.seg[si] = a;
            t.nseg++;
            j.pending = 2;
            j.seg = (int)si;
            return true;
        }
        if (f.state == ST_CHILD) {
            const int n = chain_up(j, (int)j.st.size() - 1);
            if (n == 0) return false;  // no upsweep room left in this step
            j.chain = n;
            j.pending = 3;
            return true;
        }
        if (f.state == ST_UPDONE) {
            if (f.nx.b) j.top = f.mark;
            f.nx = Engine::TLevel{};
            if (f.chunking && f.off + f.m < f.end) {
                f.off += f.m;
                f.state = ST_START;
                continue;
            }
            j.st.pop_back();
            if (!j.st.empty()) j.st.back().state = ST_CHILD;
            continue;
        }
        return false;  // an action is in flight
    }
}

void Server::run() {
    {
        pf_status gs = hipSetDevice(e_.device) == hipSuccess ? PF_OK : PF_EDEVICE;
        for (int f = 0; f < 2; f++)
            for (int l = 0; l < 2; l++)
                for (int r = 0; r <= kUpRounds && gs == PF_OK; r++) gs = build_graph(f, l, r);
        std::lock_guard<std::mutex> g(mu_);
        graph_status_ = gs;
    }
    cv_.notify_all();
    if (graph_status_ != PF_OK) return;
    std::vector<int> active;
    std::vector<int> admitted;
    for (;;) {
        admitted.clear();
        if (active.empty()) idle_spin();
        {
            std::unique_lock<std::mutex> g(mu_);
            cv_.wait(g, [&] { return stop_ || !queue_.empty() || !active.empty(); });
            if (stop_ && queue_.empty() && active.empty()) break;
            // admission: a free slot and a free region of the predicted size
            const double tnow = now_s();
            size_t scanned = 0;
            for (auto it = queue_.begin(); it != queue_.end() && scanned < 4096;) {
                if (free_slots_.empty() || (int)admitted.size() >= kMaxRoots) break;
                scanned++;
                if (it->depth == 0) {
                    pf_result res{it->tag, 1, PF_OK, 0, it->t_submit, tnow};
                    done_.push_back(res);
                    it = queue_.erase(it);
                    continue;
                }
                u64 off = 0;
                if (!region_alloc(it->want, &off)) {
                    // an old request that does not fit stops younger ones from passing it
                    if (tnow - it->t_submit > 0.050) break;
                    ++it;
                    continue;
                }
                Job jb = std::move(*it);
                it = queue_.erase(it);
                jb.base = off;
                jb.cap = jb.want;
                const int slot = free_slots_.back();
                free_slots_.pop_back();
                jb.slot = slot;
                jobs_[(size_t)slot] = std::move(jb);
                admitted.push_back(slot);
            }
            queue_idle_ = queue_.empty();
            queued_.store(queue_.size(), std::memory_order_release);
        }
        StepTable& t = *h_tab_;
        t.nseg = t.nup = t.nroot = 0;
        t.tile_ctr = 0;
        t.nlate = 0;
        t.tile_ctr3 = 0;
        late_ok_ = queue_idle_ && active.size() + admitted.size() <= kLateMaxJobs;
        for (auto& v : up_) v.clear();
        nup_planned_ = 0;
        // running jobs first, then the new ones: each new root is probed and
        // counted by the start kernel, and its first level is expanded by the
        // step kernel of the same step
        for (int slot : active) {
            Job& j = jobs_[(size_t)slot];
            if (j.pending == 0 && !j.done) plan(j);
        }
        for (int slot : admitted) {
            Job& j = jobs_[(size_t)slot];
            const int black = pos_black_to_move(j.pos);
            RootSeg& rs = t.root[t.nroot++];
            rs.b = to_board(j.pos);
            rs.k1 = zobrist_full(e_.h_z.k[0], rs.b, black);
            rs.k2 = zobrist_full(e_.h_z.k[1], rs.b, black);
            rs.word = (u32)slot | ((u32)black << 31);
            rs.depth = j.depth;
            rs.dst_b = root_b_ + slot;
            rs.dst_k1 = root_k1_ + slot;
            rs.dst_k2 = root_k2_ + slot;
            rs.dst_w = root_w_ + slot;
            rs.acc = slot_acc_ + slot;
            j.top = 0;
            j.st.clear();
            j.pending = 1;
            if (j.depth >= 2) {
                Frame f = {};
                f.lv.b = root_b_ + slot;
                f.lv.root = root_w_ + slot;
                f.lv.k1 = root_k1_ + slot;
                f.lv.k2 = root_k2_ + slot;
                f.lv.parent = root_p_ + slot;
                f.lv.acc = slot_acc_ + slot;
                f.off = 0;
                f.m = 1;
                f.n_next = (u64)count_moves(rs.b, KSBackend{});  // what the start kernel will store
                f.L = 0;
                f.state = ST_START;
                j.st.push_back(f);
                if (!plan(j)) j.pending = 4;  // root only this step; planned with the running jobs next
            }
            active.push_back(slot);
        }
        // drop finished jobs from the active list
        active.erase(std::remove_if(active.begin(), active.end(), [&](int s) { return jobs_[(size_t)s].done; }),
                     active.end());
        if (t.nseg == 0 && nup_planned_ == 0 && t.nroot == 0) {
            // nothing can move: a job waiting for memory inside its own region
            // will never get it, so fail such jobs instead of spinning
            for (int slot : active)
                if (!jobs_[(size_t)slot].done) finish(jobs_[(size_t)slot], 0, PF_ENOMEM);
            active.clear();
            continue;
        }
        // write segments first, then leaf segments (two tile queues of one
        // launch); a segment keeps its outcome slot (its out pointers), so
        // jobs find their outcome where they planned it
        {
            u32 nw = 0;
            for (u32 i = 0; i < t.nseg; i++)
                if (t.seg[i].mode == kWrite) {
                    pos_of_[i] = nw;
                    seg_tmp_[nw++] = t.seg[i];
                }
            t.nseg_write = nw;
            for (u32 i = 0; i < t.nseg; i++)
                if (t.seg[i].mode != kWrite) {
                    pos_of_[i] = nw;
                    seg_tmp_[nw++] = t.seg[i];
                }
            std::copy(seg_tmp_.begin(), seg_tmp_.begin() + t.nseg, t.seg);
            t.tile_ctr2 = 0;
            for (u32 k = 0; k < t.nlate; k++) t.late_pos[k] = pos_of_[late_src_[k]];
        }
        // tile prefix over expand segments; a step whose levels are small
        // tiles them finer (parents per tile: the fewest from kFineMinParents
        // up that give every tile a workgroup of the mixed launch), since its
        // time is then one tile's chain of dependent steps
        u32 tp = kTTTileParents;
        for (u32 c = kFineMinParents; c < (u32)kTTTileParents; c *= 2) {
            u64 n = 0;
            for (u32 i = 0; i < t.nseg; i++) n += (t.seg[i].n_in + c - 1) / c;
            if (n <= (u64)grid_) {
                tp = c;
                break;
            }
        }
        const int fine = tp < (u32)kTTTileParents;
        t.mix_tp = tp;
        u32 acc_t = 0;
        for (u32 i = 0; i < t.nseg; i++) {
            t.tile_prefix[i] = acc_t;
            acc_t += (t.seg[i].n_in + tp - 1) / tp;
        }
        t.tile_prefix[t.nseg] = acc_t;
        t.mix_write_wgs = mix_write_wgs(t.tile_prefix[t.nseg_write], acc_t - t.tile_prefix[t.nseg_write]);
        // upsweep segments laid out round by round
        u32 acc_u = 0;
        for (int r = 0; r < kUpRounds; r++) {
            t.up_round[r] = t.nup;
            for (const UpSeg& u : up_[(size_t)r]) {
                t.up[t.nup] = u;
                t.up_prefix[t.nup] = acc_u;
                acc_u += u.n + u.ndup;  // nodes, then duplicates
                t.nup++;
            }
        }
        t.up_round[kUpRounds] = t.nup;
        t.up_prefix[t.nup] = acc_u;
        const double t0 = now_s();
        int rounds = 0;  // rounds are filled in order: round r + 1 only after round r
        while (rounds < kUpRounds && t.up_round[rounds + 1] > t.up_round[rounds]) rounds++;
        t.seq = ++seq_;
        const bool ok =
            hipGraphLaunch(exec_[fine][t.nlate > 0][rounds], e_.stream) == hipSuccess && wait_word(&h_tab_->done, t.seq);
        const double t1 = now_s();
        {
            std::lock_guard<std::mutex> g(mu_);
            steps_++;
            segs_ += t.nseg + t.nup + t.nroot;
            busy_ += t1 - t0;
            if (ok && e_.d_tt) {
                probes_ = h_tab_->stats[0];
                hits_ = h_tab_->stats[1];
                for (u32 i = 0; i < t.nroot; i++) {
                    if (t.root[i].depth < kMinTTDepth) continue;
                    root_probes_++;
                    root_hits_ += t.rout[i].hit;
                }
            }
        }
        if (!ok) {
            for (int slot : active)
                if (!jobs_[(size_t)slot].done) finish(jobs_[(size_t)slot], 0, PF_EDEVICE);
            active.clear();
            continue;
        }
        // outcomes
        u64 step_dups = 0;
        for (int slot : active) {
            Job& j = jobs_[(size_t)slot];
            if (j.done) continue;
            const int pend = j.pending;
            j.pending = 0;
            if (pend == 1 || pend == 4) {  // a new root alone this step
                if (j.depth == 1) finish(j, h_tab_->acc[j.slot], PF_OK);
                continue;
            }
            if (pend == 3) {  // upsweeps alone
                if (after_chain(j, j.chain)) finish(j, h_tab_->acc[j.slot], PF_OK);
                continue;
            }
            if (pend != 2) continue;
            const SegOut& o = t.out[j.seg];
            if (o.err) {
                finish(j, 0, PF_ENOMEM);
                continue;
            }
            step_dups += o.ndup;
            Frame& f = j.st.back();
            if (f.state == ST_WRITE_LATE) {  // written, counted and upswept in this step
                if (after_chain(j, j.chain)) finish(j, h_tab_->acc[j.slot], PF_OK);
                continue;
            }
            if (f.state == ST_LEAF) {
                j.st.pop_back();
                if (j.st.empty() || after_chain(j, j.chain)) finish(j, h_tab_->acc[j.slot], PF_OK);
                continue;
            }
            f.nx.ndup = o.ndup;
            if (o.out_count == 0) {  // every child was answered by the table
                f.written = 0;
                f.state = ST_UPDONE;
            } else {  // descend into the written children
                f.written = o.out_count;
                f.state = ST_WAITCHILD;
                Frame c = {};
                c.lv = f.nx;
                c.off = 0;
                c.m = o.out_count;
                c.n_next = o.total;
                c.L = f.L + 1;
                c.state = ST_START;
                j.st.push_back(c);
            }
        }
        active.erase(std::remove_if(active.begin(), active.end(), [&](int s) { return jobs_[(size_t)s].done; }),
                     active.end());
        if (step_dups) {
            std::lock_guard<std::mutex> g(mu_);
            dups_ += step_dups;
        }
    }
    // the last step's outcomes were seen in the host table; its graph is
    // complete on the stream too before the server's memory is freed
    hipStreamSynchronize(e_.stream);
}

}  // namespace pf
