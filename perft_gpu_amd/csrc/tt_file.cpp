// tt_file.cpp -- saving and restoring the transposition table.
//
// The reference keeps deep subtree counts across runs in a disk hash
// (diskTTProbe / diskTTStore on perfthash.dat, launcher.h:700-789): a
// chained file table probed and written entry by entry under a lock file.
// Here the table lives in HBM during a run and the file is a snapshot of it:
// streamed out after a run and back in before the next, through two pinned
// staging buffers so that disk and PCIe transfers overlap.
//
// File: a 64-byte header, then buckets x 32 bytes exactly as in HBM (two
// 16-byte entries {w0 = k2 ^ w1, w1 = depth << 56 | count}).  An entry carries
// its own check (w0 ^ w1 must equal the probing position's second key), so a
// damaged entry can only miss; the header pins what makes entries meaningful:
// the entry encoding and the Zobrist keys.
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>

#include "engine.h"
#include "tt_kernels.h"

namespace pf {

namespace {

constexpr char kMagic[8] = {'P', 'F', 'T', 'T', 'A', 'B', 'L', 'E'};
constexpr u32 kFormat = 1;             // entry encoding and key scheme
constexpr u64 kBucketBytes = 32;
constexpr u64 kChunkBytes = 64ULL << 20;  // staging chunk (a multiple of a bucket)

struct TTFileHeader {
    char magic[8];
    u32 format;
    u32 bucket_bytes;
    u64 buckets;       // a power of two
    u64 zobrist_fp;    // fingerprint of both Zobrist key tables
    u64 reserved[4];
};
static_assert(sizeof(TTFileHeader) == 64, "table file header is 64 bytes");

// FNV-1a over the 64-bit words of both key tables, with an extra shift-xor
// per word so that every bit of a word reaches the result.
u64 zobrist_fingerprint(const ZTables& z) {
    const u64* w = reinterpret_cast<const u64*>(&z);
    u64 h = 0xcbf29ce484222325ULL;
    for (size_t i = 0; i < sizeof(ZTables) / 8; i++) {
        h ^= w[i];
        h *= 0x100000001b3ULL;
        h ^= h >> 29;
    }
    return h;
}

// Two pinned host buffers, two device buffers (loading only) and their
// completion events; released on every path out.
struct Staging {
    char* host[2] = {nullptr, nullptr};
    uint4* dev[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    FILE* f = nullptr;
    hipStream_t stream;
    explicit Staging(hipStream_t s) : stream(s) {}
    ~Staging() {
        hipStreamSynchronize(stream);  // nothing in flight reads or writes the buffers
        for (int i = 0; i < 2; i++) {
            if (done[i]) { hipEventSynchronize(done[i]); hipEventDestroy(done[i]); }
            if (host[i]) hipHostFree(host[i]);
            if (dev[i]) hipFree(dev[i]);
        }
        if (f) fclose(f);
    }
    bool init(u64 bytes, bool device) {
        for (int i = 0; i < 2; i++) {
            if (hipHostMalloc(reinterpret_cast<void**>(&host[i]), bytes, hipHostMallocDefault) != hipSuccess) {
                host[i] = nullptr;
                return false;
            }
            if (device && hipMalloc(reinterpret_cast<void**>(&dev[i]), bytes) != hipSuccess) {
                dev[i] = nullptr;
                return false;
            }
            if (hipEventCreateWithFlags(&done[i], hipEventDisableTiming) != hipSuccess) {
                done[i] = nullptr;
                return false;
            }
        }
        return true;
    }
};

}  // namespace

pf_status Engine::tt_save(const char* path) {
    if (!d_tt || !path || tt_split) return PF_EINVAL;  // the file holds the one-table layout
    if (hipSetDevice(device) != hipSuccess) return PF_EDEVICE;
    Staging g(stream);
    const u64 total = tt_buckets * kBucketBytes, chunk = std::min(total, kChunkBytes);
    if (!g.init(chunk, false)) return PF_ENOMEM;
    g.f = fopen(path, "wb");
    if (!g.f) return PF_EIO;
    TTFileHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, kMagic, 8);
    h.format = kFormat;
    h.bucket_bytes = (u32)kBucketBytes;
    h.buckets = tt_buckets;
    h.zobrist_fp = zobrist_fingerprint(h_z);
    if (fwrite(&h, sizeof(h), 1, g.f) != 1) return PF_EIO;
    // chunk c goes to host[c & 1]; while chunk c is written out, chunk c + 1
    // is already on its way (its buffer held chunk c - 1, written before)
    const u64 nchunks = (total + chunk - 1) / chunk;
    const char* src = reinterpret_cast<const char*>(d_tt);
    auto issue = [&](u64 c) -> bool {
        const u64 off = c * chunk, len = std::min(chunk, total - off);
        return hipMemcpyAsync(g.host[c & 1], src + off, len, hipMemcpyDeviceToHost, stream) == hipSuccess &&
               hipEventRecord(g.done[c & 1], stream) == hipSuccess;
    };
    if (!issue(0)) return PF_EDEVICE;
    for (u64 c = 0; c < nchunks; c++) {
        if (c + 1 < nchunks && !issue(c + 1)) return PF_EDEVICE;
        if (hipEventSynchronize(g.done[c & 1]) != hipSuccess) return PF_EDEVICE;
        const u64 len = std::min(chunk, total - c * chunk);
        if (fwrite(g.host[c & 1], 1, len, g.f) != len) return PF_EIO;
    }
    FILE* f = g.f;
    g.f = nullptr;
    return fclose(f) == 0 ? PF_OK : PF_EIO;
}

pf_status Engine::tt_load(const char* path) {
    if (!d_tt || !path || tt_split) return PF_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return PF_EDEVICE;
    Staging g(stream);
    g.f = fopen(path, "rb");
    if (!g.f) return PF_EIO;
    TTFileHeader h;
    if (fread(&h, sizeof(h), 1, g.f) != 1) return PF_EINVAL;
    if (memcmp(h.magic, kMagic, 8) || h.format != kFormat || h.bucket_bytes != kBucketBytes || h.buckets < 1024 ||
        (h.buckets & (h.buckets - 1)) || h.buckets > (1ULL << 40))
        return PF_EINVAL;
    if (h.zobrist_fp != zobrist_fingerprint(h_z)) return PF_EINVAL;
    struct stat sb;
    if (fstat(fileno(g.f), &sb) != 0) return PF_EIO;
    const u64 total = h.buckets * kBucketBytes;
    if ((u64)sb.st_size != sizeof(h) + total) return PF_EINVAL;
    const u64 chunk = std::min(total, kChunkBytes), nchunks = (total + chunk - 1) / chunk;
    if (!g.init(chunk, true)) return PF_ENOMEM;
    pf_status s = tt_clear();
    if (s != PF_OK) return s;
    // chunk c: file -> host[c & 1] -> dev[c & 1] -> refit into the table;
    // the event on slot c & 1 covers chunk c - 2's copy and kernel
    for (u64 c = 0; c < nchunks; c++) {
        const int k = (int)(c & 1);
        if (c >= 2 && hipEventSynchronize(g.done[k]) != hipSuccess) return PF_EDEVICE;
        const u64 len = std::min(chunk, total - c * chunk);
        if (fread(g.host[k], 1, len, g.f) != len) return PF_EIO;
        if (hipMemcpyAsync(g.dev[k], g.host[k], len, hipMemcpyHostToDevice, stream) != hipSuccess ||
            launch_tt_refit(g.dev[k], c * chunk / kBucketBytes, len / kBucketBytes, h.buckets, d_tt, tt_buckets - 1,
                            n_cu, stream) != hipSuccess ||
            hipEventRecord(g.done[k], stream) != hipSuccess)
            return PF_EDEVICE;
    }
    return hipStreamSynchronize(stream) == hipSuccess ? PF_OK : PF_EDEVICE;
}

}  // namespace pf
