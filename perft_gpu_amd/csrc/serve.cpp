// serve.cpp -- continuous-batching perft server (no reference counterpart: the
// reference answers one perft per process run, perft.cu:133-137).
//
// Requests (position, depth) are admitted while a job slot is free; each slot
// owns an equal share of the frontier arena, so admission is by frontier
// memory and a job never waits on another job's memory.  A job is the
// depth-first chunked level driver of engine_tt.cpp rewritten as an explicit
// frame stack, so that it can advance one GPU action at a time.
//
// One scheduler step = one replay of a captured graph:
//   H2D step table -> roots kernel (new requests: table probe at full depth,
//   move count) -> step kernel (the expand/leaf levels of every active job in
//   one persistent launch over a shared tile queue) -> upsweep kernel (the
//   finished levels of every job) -> D2H of per-segment outcomes and of the
//   slot accumulators.
// Requests that arrive while a step runs join the next step; their work
// overlaps on the device with everyone else's.  All jobs share the lockless
// transposition table.
#include "serve.h"

#include <string.h>

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
enum { ST_START = 0, ST_WRITE, ST_LEAF, ST_WAITCHILD, ST_CHILD, ST_UP, ST_UPDONE };

Server::Server(Engine& e, int max_jobs, u64 job_bytes) : e_(e), max_jobs_(max_jobs), job_bytes_(job_bytes) {}

Server::~Server() { stop(); }

pf_status Server::start() {
    if (running_) return PF_OK;
    if (max_jobs_ < 1 || max_jobs_ > kMaxRoots) return PF_EINVAL;
    if (!job_bytes_) job_bytes_ = e_.arena_bytes / (u64)max_jobs_;
    job_bytes_ &= ~(kAl - 1);
    if (job_bytes_ * (u64)max_jobs_ > e_.arena_bytes || job_bytes_ < (1u << 20)) return PF_ENOMEM;
    if (hipSetDevice(e_.device) != hipSuccess) return PF_EDEVICE;
    if (hipHostMalloc(&h_tab_, sizeof(StepTable), hipHostMallocDefault) != hipSuccess) return PF_ENOMEM;
    if (hipHostMalloc(&h_acc_, sizeof(u64) * max_jobs_, hipHostMallocDefault) != hipSuccess) return PF_ENOMEM;
    if (hipMalloc(&d_tab_, sizeof(StepTable)) != hipSuccess) return PF_ENOMEM;
    const int n = max_jobs_;
    if (hipMalloc(&root_b_, sizeof(Board) * n) != hipSuccess || hipMalloc(&root_w_, 4 * n) != hipSuccess ||
        hipMalloc(&root_k1_, 8 * n) != hipSuccess || hipMalloc(&root_k2_, 8 * n) != hipSuccess ||
        hipMalloc(&root_p_, 4 * n) != hipSuccess || hipMalloc(&slot_acc_, 8 * n) != hipSuccess)
        return PF_ENOMEM;
    hipMemset(root_p_, 0, 4 * n);
    memset(h_tab_, 0, sizeof(StepTable));
    jobs_.assign((size_t)n, Job{});
    free_slots_.clear();
    for (int i = n - 1; i >= 0; i--) free_slots_.push_back(i);
    grid_ = step_grid(e_.backend, e_.n_cu);
    up_grid_ = e_.n_cu * 4;
    pf_status s = build_graph();
    if (s != PF_OK) return s;
    stop_ = false;
    running_ = true;
    th_ = std::thread([this] { run(); });
    return PF_OK;
}

pf_status Server::build_graph() {
    hipStream_t st = e_.stream;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return PF_EDEVICE;
    hipMemcpyAsync(d_tab_, h_tab_, sizeof(StepTable), hipMemcpyHostToDevice, st);
    launch_step(e_.backend, e_.d_tab, e_.d_z, d_tab_, e_.d_tt, e_.tt_buckets ? e_.tt_buckets - 1 : 0, grid_, up_grid_,
                st);
    hipMemcpyAsync(&h_tab_->out[0], &d_tab_->out[0], sizeof(h_tab_->out) + sizeof(h_tab_->rout),
                   hipMemcpyDeviceToHost, st);
    hipMemcpyAsync(h_acc_, slot_acc_, sizeof(u64) * max_jobs_, hipMemcpyDeviceToHost, st);
    if (hipStreamEndCapture(st, &graph_) != hipSuccess) return PF_EDEVICE;
    if (hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0) != hipSuccess) return PF_EDEVICE;
    return PF_OK;
}

void Server::stop() {
    if (running_) {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
        running_ = false;
    }
    if (exec_) hipGraphExecDestroy(exec_);
    if (graph_) hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    hipFree(d_tab_);
    hipFree(root_b_);
    hipFree(root_w_);
    hipFree(root_k1_);
    hipFree(root_k2_);
    hipFree(root_p_);
    hipFree(slot_acc_);
    hipHostFree(h_tab_);
    hipHostFree(h_acc_);
    d_tab_ = nullptr; root_b_ = nullptr; root_w_ = nullptr; root_k1_ = nullptr; root_k2_ = nullptr;
    root_p_ = nullptr; slot_acc_ = nullptr; h_tab_ = nullptr; h_acc_ = nullptr;
}

pf_status Server::submit(const pf_pos48& pos, int depth, u64 tag) {
    if (depth < 0 || depth > 16) return PF_EINVAL;
    Job j;
    j.tag = tag;
    j.depth = depth;
    j.pos = pos;
    j.t_submit = now_s();
    {
        std::lock_guard<std::mutex> g(mu_);
        if (!running_ || stop_) return PF_EBUSY;
        if (queue_.size() > 1000000) return PF_EBUSY;
        queue_.push_back(std::move(j));
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

void Server::counters(u64* steps, u64* segs, double* busy) {
    std::lock_guard<std::mutex> g(mu_);
    *steps = steps_;
    *segs = segs_;
    *busy = busy_;
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
    j.slot = -1;
}

bool Server::alloc_level(Job& j, u64 n, Engine::TLevel* lv) {
    char* base = e_.arena + (u64)j.slot * job_bytes_;
    auto take = [&](u64 bytes) -> char* {
        bytes = (bytes + kAl - 1) & ~(kAl - 1);
        if (j.top + bytes > job_bytes_) return nullptr;
        char* p = base + j.top;
        j.top += bytes;
        return p;
    };
    lv->b = reinterpret_cast<Board*>(take(n * sizeof(Board)));
    lv->root = reinterpret_cast<u32*>(take(n * 4));
    lv->k1 = reinterpret_cast<u64*>(take(n * 8));
    lv->k2 = reinterpret_cast<u64*>(take(n * 8));
    lv->parent = reinterpret_cast<u32*>(take(n * 4));
    lv->acc = reinterpret_cast<unsigned long long*>(take(n * 8));
    return lv->b && lv->root && lv->k1 && lv->k2 && lv->parent && lv->acc;
}

// Advance job j to its next GPU action.  Returns true if it added a segment
// to the step table; finishes the job when its frame stack empties.
bool Server::plan(Job& j) {
    StepTable& t = *h_tab_;
    for (;;) {
        if (j.st.empty()) {
            finish(j, h_acc_[j.slot], PF_OK);
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
            a.tt_mask = e_.tt_buckets ? e_.tt_buckets - 1 : 0;
            a.child_depth = r - 1;
            a.stats = e_.d_ttstats;
            const u32 si = t.nseg;
            a.out_count = &d_tab_->out[si].out_count;
            a.err = &d_tab_->out[si].err;
            a.tile_ctr = nullptr;
            if (r == 2) {
                a.mode = kLeaf;
                f.state = ST_LEAF;
                f.nx = Engine::TLevel{};
            } else {
                const u64 free_b = job_bytes_ - j.top;
                u64 cap;
                if (!f.chunking && f.n_next * kTNode + 8 * kAl <= free_b / 2 + free_b / 4) {
                    cap = f.n_next;
                } else {
                    if (!f.chunking) {
                        f.chunking = true;
                        f.end = f.off + f.m;
                    }
                    const u64 per = 218 * kTNode;
                    const u64 pm = std::max<u64>(1, (free_b / 2) / per);
                    f.m = std::min(pm, f.end - f.off);
                    a.n_in = (u32)f.m;
                    cap = f.m * 218;
                }
                f.mark = j.top;
                if (!alloc_level(j, cap, &f.nx)) {
                    j.top = f.mark;
                    return false;  // cannot happen with the half-free rule; wait
                }
                f.cap = cap;
                a.mode = kWrite;
                a.out = f.nx.b;
                a.out_root = f.nx.root;
                a.out_k1 = f.nx.k1;
                a.out_k2 = f.nx.k2;
                a.out_parent = f.nx.parent;
                a.out_acc = f.nx.acc;
                a.out_cap = cap;
                a.total = &d_tab_->out[si].total;
                f.state = ST_WRITE;
            }
            t.seg[si] = a;
            t.nseg++;
            j.pending = 2;
            j.seg = (int)si;
            return true;
        }
        if (f.state == ST_CHILD) {
            if (t.nup >= (u32)kMaxUps) return false;
            const u32 ui = t.nup++;
            UpSeg& u = t.up[ui];
            u.parent = f.nx.parent;
            u.k1 = f.nx.k1;
            u.k2 = f.nx.k2;
            u.acc = f.nx.acc;
            u.parent_acc = f.lv.acc;
            u.n = (u32)f.written;
            u.depth = r - 1;
            f.state = ST_UP;
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
    hipSetDevice(e_.device);
    std::vector<int> active;
    std::vector<int> admitted;
    for (;;) {
        admitted.clear();
        {
            std::unique_lock<std::mutex> g(mu_);
            cv_.wait(g, [&] { return stop_ || !queue_.empty() || !active.empty(); });
            if (stop_ && queue_.empty() && active.empty()) break;
            // admission: one slot (an equal share of the frontier arena) per job
            while (!queue_.empty() && !free_slots_.empty() && (int)admitted.size() < kMaxRoots) {
                Job jb = std::move(queue_.front());
                queue_.pop_front();
                if (jb.depth == 0) {
                    pf_result res{jb.tag, 1, PF_OK, 0, jb.t_submit, now_s()};
                    done_.push_back(res);
                    continue;
                }
                const int slot = free_slots_.back();
                free_slots_.pop_back();
                jb.slot = slot;
                jobs_[(size_t)slot] = std::move(jb);
                admitted.push_back(slot);
            }
        }
        StepTable& t = *h_tab_;
        t.nseg = t.nup = t.nroot = 0;
        t.tile_ctr = 0;
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
            j.pending = 1;
            j.top = 0;
            j.st.clear();
        }
        for (int slot : active) {
            Job& j = jobs_[(size_t)slot];
            if (j.pending == 0 && !j.done) plan(j);
        }
        // drop finished jobs from the active list
        active.erase(std::remove_if(active.begin(), active.end(), [&](int s) { return jobs_[(size_t)s].done; }),
                     active.end());
        if (t.nseg == 0 && t.nup == 0 && t.nroot == 0) {
            // nothing can move: a job waiting for memory inside its own region
            // will never get it, so fail such jobs instead of spinning
            for (int slot : active)
                if (!jobs_[(size_t)slot].done) finish(jobs_[(size_t)slot], 0, PF_ENOMEM);
            active.clear();
            continue;
        }
        // tile prefix over expand segments
        u32 acc_t = 0;
        for (u32 i = 0; i < t.nseg; i++) {
            t.tile_prefix[i] = acc_t;
            acc_t += (t.seg[i].n_in + 127) / 128;
        }
        t.tile_prefix[t.nseg] = acc_t;
        u32 acc_u = 0;
        for (u32 i = 0; i < t.nup; i++) {
            t.up_prefix[i] = acc_u;
            acc_u += t.up[i].n;
        }
        t.up_prefix[t.nup] = acc_u;
        memset(t.out, 0, sizeof(SegOut) * t.nseg);
        const double t0 = now_s();
        const bool ok = hipGraphLaunch(exec_, e_.stream) == hipSuccess && hipStreamSynchronize(e_.stream) == hipSuccess;
        const double t1 = now_s();
        {
            std::lock_guard<std::mutex> g(mu_);
            steps_++;
            segs_ += t.nseg + t.nup + t.nroot;
            busy_ += t1 - t0;
        }
        if (!ok) {
            for (int slot : active) finish(jobs_[(size_t)slot], 0, PF_EDEVICE);
            for (int slot : admitted) finish(jobs_[(size_t)slot], 0, PF_EDEVICE);
            active.clear();
            continue;
        }
        // outcomes
        for (size_t i = 0; i < admitted.size(); i++) {
            Job& j = jobs_[(size_t)admitted[i]];
            j.pending = 0;
            const RootOut& ro = t.rout[i];
            if (ro.hit) { finish(j, h_acc_[j.slot], PF_OK); continue; }
            if (j.depth == 1) { finish(j, ro.count, PF_OK); continue; }
            Frame f = {};
            f.lv.b = root_b_ + j.slot;
            f.lv.root = root_w_ + j.slot;
            f.lv.k1 = root_k1_ + j.slot;
            f.lv.k2 = root_k2_ + j.slot;
            f.lv.parent = root_p_ + j.slot;
            f.lv.acc = slot_acc_ + j.slot;
            f.off = 0;
            f.m = 1;
            f.n_next = ro.count;
            f.L = 0;
            f.state = ST_START;
            j.st.push_back(f);
            active.push_back(j.slot);
        }
        for (int slot : active) {
            Job& j = jobs_[(size_t)slot];
            if (j.done || j.st.empty()) continue;
            Frame& f = j.st.back();
            if (j.pending == 2) {
                const SegOut& o = t.out[j.seg];
                j.pending = 0;
                if (o.err) { finish(j, 0, PF_ENOMEM); continue; }
                if (f.state == ST_LEAF) {
                    j.st.pop_back();
                    if (!j.st.empty()) j.st.back().state = ST_CHILD;
                } else if (o.out_count == 0) {
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
            } else if (j.pending == 3) {
                j.pending = 0;
                f.state = ST_UPDONE;
            }
        }
        active.erase(std::remove_if(active.begin(), active.end(), [&](int s) { return jobs_[(size_t)s].done; }),
                     active.end());
    }
}

}  // namespace pf
