// noalloc.hip -- the stream-ordered batch (PF_MODE_STREAM) without any
// allocation: for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops, harnesses that forbid it).
//
// The control block, the converted roots and every frontier level live in
// one fixed device array of kScratchBytes, `scratch`, instead of the
// context's arena; the per-root counts accumulate directly in the caller's
// output array.  The array is loaded with this file's code object, the first
// time a call on a device asks for its address, and each device keeps its
// own copy until the process ends.  Calls on one device share it: calls that
// overlap on one GPU (two contexts, or two threads on two contexts, of the
// same device) race.
#include <algorithm>
#include <mutex>

#include "engine.h"

namespace pf {

// 512 MiB: the largest C2-sized batch (256 positions, perft 5: about 438 MB of
// frontier at 52 B per node) fits; larger batches return PF_ENOMEM.
constexpr u64 kScratchBytes = 512ULL << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Host code never passes `scratch` itself (its host-side name is not a device
// address): each device's address comes from hipGetSymbolAddress, once per
// device, with that device current.
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

static hipError_t scratch_address(int device, void** out) {
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

// Clears the control block and the caller's counts (a fill kernel, not a
// memset, so that the whole call can be captured in a graph).
__global__ __launch_bounds__(256) void k_noalloc_clear(PersistCtl* ctl, unsigned long long* out, u32 n,
                                                       unsigned long long fill) {
    const u32 i = blockIdx.x * 256 + threadIdx.x;
    constexpr u32 kWords = sizeof(PersistCtl) / 4;
    static_assert(sizeof(PersistCtl) % 4 == 0, "control block cleared word by word");
    if (i < kWords) reinterpret_cast<u32*>(ctl)[i] = 0;
    if (i < n) out[i] = fill;
}

static constexpr u64 align256(u64 x) { return (x + 255) & ~255ULL; }

pf_status Engine::perft_noalloc(const pf_pos48* dpos, int n, int d, u64* dout) {
    if (n <= 0) return PF_OK;
    if (hipSetDevice(device) != hipSuccess) return PF_EDEVICE;
    void* base = nullptr;
    if (scratch_address(device, &base) != hipSuccess) return PF_EDEVICE;
    char* s = static_cast<char*>(base);
    // layout: control block | roots (Board) | root ids | frontier levels
    const u64 ctl_b = align256(sizeof(PersistCtl));
    const u64 roots_b = align256((u64)n * sizeof(Board)), rid_b = align256((u64)n * 4);
    if (ctl_b + roots_b + rid_b > kScratchBytes) return PF_ENOMEM;  // the roots alone do not fit
    PersistCtl* ctl = reinterpret_cast<PersistCtl*>(s);
    Board* roots = reinterpret_cast<Board*>(s + ctl_b);
    u32* rid = reinterpret_cast<u32*>(s + ctl_b + roots_b);
    unsigned long long* out = reinterpret_cast<unsigned long long*>(dout);

    const u32 words = (u32)(sizeof(PersistCtl) / 4);
    const u32 clear_n = std::max<u32>((u32)n, words);
    hipLaunchKernelGGL(k_noalloc_clear, dim3((clear_n + 255) / 256), dim3(256), 0, stream, ctl, out, (u32)n,
                       d <= 0 ? 1ULL : 0ULL);
    if (hipGetLastError() != hipSuccess) return PF_EDEVICE;
    if (d <= 0) return hipStreamSynchronize(stream) == hipSuccess ? PF_OK : PF_EDEVICE;
    if (launch_pos_to_board(dpos, (u32)n, roots, rid, stream) != hipSuccess) return PF_EDEVICE;
    PersistArgs p;
    p.roots = roots;
    p.root_ids = rid;
    p.n_roots = (u32)n;
    p.depth = d;
    p.arena = s + ctl_b + roots_b + rid_b;
    p.arena_bytes = kScratchBytes - (ctl_b + roots_b + rid_b);
    p.ctl = ctl;
    p.sums = out;
    if (timing && hipEventRecord(ev0, stream) != hipSuccess) return PF_EDEVICE;
    const int var[2] = {tune[backend].s_leaf, tune[backend].s_write};
    if (launch_stream(backend, DevTables{d_tab}, p, sgrid[backend], var, stream) != hipSuccess) return PF_EDEVICE;
    if (timing && hipEventRecord(ev1, stream) != hipSuccess) return PF_EDEVICE;
    st.kernel_launches += (u64)d + 2;  // clear, conversion, the root count and d - 1 levels
    st.levels_run += (u64)(d - 1);
    // the overflow flag comes back through the pinned buffer allocated at init
    if (hipMemcpyAsync(h_pinned, &ctl->err, sizeof(u32), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return PF_EDEVICE;
    if (timing) {
        float ms = 0;
        hipEventElapsedTime(&ms, ev0, ev1);
        st.gpu_ms += ms;
        st.last_expand_ms = ms;
        st.last_expand_launches = 1;
    }
    const u32 err = *reinterpret_cast<const u32*>(h_pinned);
    if (err & 2) return PF_ENOMEM;  // a level did not fit the scratch array
    if (err) return PF_EDEVICE;
    return PF_OK;
}

}  // namespace pf
