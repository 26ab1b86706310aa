// Device-side helpers shared by the HIP translation units of the engine:
// the HIP error check, a growable device buffer and the context (one GPU,
// its compute and copy streams, timing events).
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "ag_internal.h"

#define HIP_CHECK(x)                                                                    \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess)                                                           \
            ag::fail(AG_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) + \
                                        " at " #x);                                     \
    } while (0)

constexpr int kMaxChunks = 16;                    // pipeline stages of a batch with host outputs (at most)
constexpr int kBatchChunks = 10;                  // chunks a batch with host outputs is cut into
constexpr int kBatchTaper = 3;                    // of which the last ones halve in size

// Device blocks that destroyed plans released, kept by their context for
// the next plan: a run of plans (one per receptor, or one per batch) then
// allocates from HIP once.  Best fit among blocks at most twice the request
// (plus 1 MiB); at most kMaxHeld bytes are held, the rest is freed.
//
// Arena form (DevPool(base, size)): a fixed device block the pool does not
// own, handed out front to back in 256-byte steps and never returned or
// freed; a request past its end fails with AG_ERR_UNSUPPORTED.  Buffers
// drawing on it therefore never call hipMalloc or hipFree.
struct DevPool {
    static constexpr size_t kMaxHeld = size_t(16) << 30;
    std::mutex mu;
    std::vector<std::pair<size_t, void *>> blocks;   // (capacity, pointer)
    size_t held = 0;
    char *arena = nullptr;
    size_t arena_size = 0, arena_used = 0;
    DevPool() = default;
    DevPool(void *base, size_t size) : arena(static_cast<char *>(base)), arena_size(size) {}
    void *take(size_t bytes, size_t &cap) {
        std::lock_guard<std::mutex> lk(mu);
        if (arena) {
            const size_t want = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
            if (want > arena_size - arena_used)
                ag::fail(AG_ERR_UNSUPPORTED, "the map set needs more device memory than the " +
                                                 std::to_string(arena_size >> 20) +
                                                 " MiB scratch of the allocation-free entry point holds");
            void *p = arena + arena_used;
            arena_used += want;
            cap = want;
            return p;
        }
        int best = -1;
        for (int i = 0; i < (int)blocks.size(); i++) {
            const size_t c = blocks[i].first;
            if (c >= bytes && c <= 2 * bytes + (size_t(1) << 20) && (best < 0 || c < blocks[best].first)) best = i;
        }
        if (best < 0) return nullptr;
        void *p = blocks[best].second;
        cap = blocks[best].first;
        held -= cap;
        blocks.erase(blocks.begin() + best);
        return p;
    }
    void give(void *p, size_t cap) {
        std::lock_guard<std::mutex> lk(mu);
        if (arena) return;
        if (held + cap > kMaxHeld) {
            (void)hipFree(p);
            return;
        }
        blocks.emplace_back(cap, p);
        held += cap;
    }
    ~DevPool() {
        for (auto &b : blocks) (void)hipFree(b.second);
    }
};

// Growable device buffer.  Capacities are rounded up to 1/8 of the next
// power of two, so buffers of similar size can reuse each other's blocks.
struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    std::shared_ptr<DevPool> pool;
    static size_t rounded(size_t bytes) {
        size_t p2 = 256;
        while (p2 < bytes) p2 <<= 1;
        const size_t step = std::max<size_t>(256, p2 / 8);
        return (bytes + step - 1) / step * step;
    }
    void release() {
        if (!ptr) return;
        if (pool) pool->give(ptr, cap);
        else (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    void *get(size_t bytes) {
        if (bytes == 0) bytes = 16;
        if (bytes > cap) {
            release();
            if (pool) ptr = pool->take(bytes, cap);
            if (!ptr) {
                const size_t want = rounded(bytes);
                if (hipMalloc(&ptr, want) != hipSuccess) {
                    (void)hipGetLastError();
                    ptr = nullptr;
                    ag::fail(AG_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
                }
                cap = want;
            }
        }
        return ptr;
    }
    DevBuf() = default;
    explicit DevBuf(std::shared_ptr<DevPool> p) : pool(std::move(p)) {}
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
};

// Grow-only pinned host buffer: staging for uploads, so they run as DMA
// from page-locked memory.
struct HostStage {
    void *ptr = nullptr;
    size_t cap = 0;
    void *get(size_t bytes) {
        if (bytes > cap) {
            if (ptr) (void)hipHostFree(ptr);
            ptr = nullptr;
            cap = 0;
            const size_t want = DevBuf::rounded(bytes);
            if (hipHostMalloc(&ptr, want, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                ptr = nullptr;
                ag::fail(AG_ERR_OOM, "pinned host allocation of " + std::to_string(bytes) + " bytes failed");
            }
            cap = want;
        }
        return ptr;
    }
    HostStage() = default;
    HostStage(const HostStage &) = delete;
    HostStage &operator=(const HostStage &) = delete;
    ~HostStage() {
        if (ptr) (void)hipHostFree(ptr);
    }
};

struct ag_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;   // device-to-host copies overlapping compute
    hipStream_t stream2 = nullptr;       // second compute stream: alternate chunks of a batch
    hipEvent_t cells_ev = nullptr;       // cell list done (orders stream2 after it)
    hipEvent_t near_ev = nullptr;        // near field of a map set done
    hipEvent_t far_ev = nullptr;         // first far-field graph of a batch done
    hipEvent_t far2_ev = nullptr;        // second far-field graph of a batch done (on stream2)
    hipEvent_t chunk_ev[kMaxChunks + 1] = {};   // ordering between the two streams
    std::shared_ptr<DevPool> pool = std::make_shared<DevPool>();   // blocks of destroyed plans
    DevBuf maps;                       // output staging of ag_maps_plain
    HostStage stage;                   // receptor arrays of a plan on their way up
    HostStage status_word;             // pinned target of a plan's status flag after each map set
    void *scratch = nullptr;           // this device's allocation-free scratch (ag_maps_plain_noalloc)
    hipEvent_t ev[4] = {};
    float last_ms[4] = {0, 0, 0, 0};   // cell list, near, far, total of the last execute
    int n_cu = 0;
};
