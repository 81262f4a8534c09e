// Pinned staging ring for stream-ordered uploads of host-built tables.
#include <string.h>

#include <mutex>
#include <vector>

#include "halma_common.h"
#include "staging.h"

namespace halma {

namespace {

struct Slot {
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
    int device = -1;
};

constexpr int kSlots = 8;
std::mutex g_mu;
Slot g_slots[kSlots];
int g_next = 0;

}  // namespace

int upload_async(void* dst, const void* src, size_t bytes, hipStream_t stream) {
    const UploadPiece piece{0, src, bytes};
    return upload_pieces(dst, bytes, &piece, 1, stream);
}

int upload_pieces(void* dst, size_t bytes, const UploadPiece* pieces, int n_pieces,
                  hipStream_t stream) {
    if (bytes == 0) return 0;
    for (int q = 0; q < n_pieces; ++q)
        if (pieces[q].offset + pieces[q].bytes > bytes)
            return set_error(HALMA_E_ARG, "upload piece outside the region");
    int dev = 0;
    HALMA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_mu);
    Slot& s = g_slots[g_next];
    g_next = (g_next + 1) % kSlots;
    if (s.pending) {
        HALMA_HIP(hipEventSynchronize(s.done));
        s.pending = false;
    }
    if (s.done && s.device != dev) {
        HALMA_HIP(hipEventDestroy(s.done));
        s.done = nullptr;
    }
    if (!s.done) {
        HALMA_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
        s.device = dev;
    }
    if (s.cap < bytes) {
        if (s.host) HALMA_HIP(hipHostFree(s.host));
        s.host = nullptr;
        size_t cap = bytes < (1u << 16) ? (1u << 16) : bytes * 2;
        HALMA_HIP(hipHostMalloc(&s.host, cap, hipHostMallocDefault));
        s.cap = cap;
    }
    for (int q = 0; q < n_pieces; ++q)
        if (pieces[q].bytes)
            memcpy(static_cast<char*>(s.host) + pieces[q].offset, pieces[q].src, pieces[q].bytes);
    HALMA_HIP(hipMemcpyAsync(dst, s.host, bytes, hipMemcpyHostToDevice, stream));
    HALMA_HIP(hipEventRecord(s.done, stream));
    s.pending = true;
    return 0;
}

long long* pinned_words(int count) {
    static thread_local std::vector<std::pair<int, long long*>> per_dev;
    static thread_local std::vector<int> caps;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    for (size_t i = 0; i < per_dev.size(); ++i) {
        if (per_dev[i].first == dev) {
            if (caps[i] >= count) return per_dev[i].second;
            (void)hipHostFree(per_dev[i].second);
            void* p = nullptr;
            if (hipHostMalloc(&p, sizeof(long long) * count, hipHostMallocDefault) !=
                hipSuccess)
                return nullptr;
            per_dev[i].second = static_cast<long long*>(p);
            caps[i] = count;
            return per_dev[i].second;
        }
    }
    void* p = nullptr;
    int cap = count < 64 ? 64 : count;
    if (hipHostMalloc(&p, sizeof(long long) * cap, hipHostMallocDefault) != hipSuccess)
        return nullptr;
    per_dev.emplace_back(dev, static_cast<long long*>(p));
    caps.push_back(cap);
    return per_dev.back().second;
}

}  // namespace halma
