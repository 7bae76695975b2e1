// serve.h -- continuous-batching perft server.
#pragma once
#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.h"

namespace pf {

double now_s();  // library clock, seconds

class Server {
  public:
    Server(Engine& e, int max_jobs, u64 job_bytes);
    ~Server();
    pf_status start();
    void stop();
    pf_status submit(const pf_pos48& pos, int depth, u64 tag);
    int poll(pf_result* out, int max);
    // probes / hits: table probes and hits of the expand segments and of the
    // roots (remaining depth >= 3) since the server started
    void counters(u64* steps, u64* segs, double* busy_s, u64* dups, u64* probes, u64* hits);

  private:
    struct Frame {
        Engine::TLevel lv;
        u64 off, m, n_next, end;  // parents [off, off+m) of [.., end) when chunking
        int L;
        int state;
        bool chunking;
        Engine::TLevel nx;
        u64 cap, mark, written;
        std::vector<u32> hbsum;  // host copy of lv.bsum, fetched when the level first chunks
    };
    struct Job {
        u64 tag = 0;
        int depth = 0;
        double t_submit = 0;
        int slot = -1;
        pf_pos48 pos{};
        u64 base = 0, cap = 0;  // the job's arena region [base, base + cap)
        u64 top = 0;  // stack top within the job's arena region
        u64 want = 0;  // predicted frontier bytes (the region it asks for)
        std::vector<Frame> st;
        int pending = 0;  // 0 none, 1/4 new root only, 2 expand (+ upsweep chain), 3 upsweep chain
        int chain = 0;    // upsweeps chained in the pending action
        int seg = -1;
        bool done = false;
    };

    void run();
    u64 predict_bytes(const pf_pos48& pos, int depth) const;
    bool region_alloc(u64 bytes, u64* off);
    void region_free(u64 off, u64 bytes);
    bool grow_region(Job& j, u64 extra);
    bool plan(Job& j);
    int chain_up(Job& j, int k, int dev_seg = -1);
    bool chunk(Job& j, Frame& f, u64 budget, u64* cap);
    bool after_chain(Job& j, int n);
    void finish(Job& j, u64 nodes, int status);
    char* take(Job& j, u64 bytes);
    bool alloc_level(Job& j, u64 n, Engine::TLevel* lv);
    pf_status build_graph(int fine, int late, int rounds);
    bool wait_word(const u32* w, u32 v);
    void idle_spin();
    u32 mix_write_wgs(u32 write_tiles, u32 leaf_tiles) const;

    Engine& e_;
    int max_jobs_;
    u64 job_bytes_;
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_;
    bool stop_ = false, running_ = false;
    // read without the lock by the idle serving thread's spin: requests
    // waiting for admission, and stop() having been called
    std::atomic<size_t> queued_{0};
    std::atomic<bool> spin_stop_{false};
    bool queue_idle_ = false;  // nothing was waiting for admission this step (server thread)
    int graph_status_ = 0;  // the serving thread's graph capture: 1 pending, else a pf_status
    std::deque<Job> queue_;        // submitted, not yet admitted
    std::vector<pf_result> done_;  // finished, not yet polled
    std::vector<Job> jobs_;        // indexed by slot
    std::vector<int> free_slots_;
    std::vector<UpSeg> up_[kUpRounds];  // this step's upsweeps, by round
    u32 nup_planned_ = 0;
    std::map<u64, u64> free_mem_;  // arena offset -> bytes of each free block (coalesced)
    u64 max_region_ = 0;
    // device / pinned resources
    StepTable* h_tab_ = nullptr;  // coherent host memory, read and written by the step's first and last kernels
    StepTable* h_dev_ = nullptr;  // h_tab_ as the device addresses it
    StepTable* d_tab_ = nullptr;
    Board* root_b_ = nullptr;
    u32* root_w_ = nullptr;
    u64* root_k1_ = nullptr;
    u64* root_k2_ = nullptr;
    u32* root_p_ = nullptr;
    unsigned long long* slot_acc_ = nullptr;  // d_tab_->acc
    u64 probes_ = 0, hits_ = 0, root_probes_ = 0, root_hits_ = 0;
    int grid_ = 0, up_grid_ = 0;
    std::vector<TTArgs> seg_tmp_;  // reordering buffer of a step's segments
    std::vector<u32> pos_of_;      // a segment's position after the reordering
    bool late_ok_ = false;         // this step may hold late leaf passes
    u32 seq_ = 0;                  // steps launched
    u32 rb_seq_ = 0;               // block-sum read-backs launched
    u32* h_rb_ = nullptr;          // pinned staging of block-sum read-backs (rb_cap_ entries)
    u64 rb_cap_ = 0;
    u32 late_src_[kMaxLate] = {};  // their write segments (indices before the reordering)
    // by tiling (1: fine), late leaf launch (1: recorded), then upsweep rounds
    hipGraph_t graph_[2][2][kUpRounds + 1] = {};
    hipGraphExec_t exec_[2][2][kUpRounds + 1] = {};
    u64 steps_ = 0, segs_ = 0, dups_ = 0;
    double busy_ = 0;
};

}  // namespace pf
