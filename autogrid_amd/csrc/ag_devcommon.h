// Device-side helpers shared by the HIP translation units of the engine:
// the HIP error check, a growable device buffer and the context (one GPU,
// its compute and copy streams, timing events).
#pragma once

#include <hip/hip_runtime.h>
#include <string>

#include "ag_internal.h"

#define HIP_CHECK(x)                                                                    \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess)                                                           \
            ag::fail(AG_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) + \
                                        " at " #x);                                     \
    } while (0)

constexpr int kMaxChunks = 16;                    // pipeline stages of a batch with host outputs

struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    void *get(size_t bytes) {
        if (bytes == 0) bytes = 16;
        if (bytes > cap) {
            if (ptr) HIP_CHECK(hipFree(ptr));
            ptr = nullptr;
            cap = 0;
            if (hipMalloc(&ptr, bytes) != hipSuccess) {
                ptr = nullptr;
                ag::fail(AG_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
            }
            cap = bytes;
        }
        return ptr;
    }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (ptr) (void)hipFree(ptr);
    }
};

struct ag_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;   // device-to-host copies overlapping compute
    hipEvent_t chunk_ev[kMaxChunks + 1] = {};   // ordering between the two streams
    DevBuf maps;                       // output staging of ag_maps_plain
    hipEvent_t ev[4] = {};
    float last_ms[4] = {0, 0, 0, 0};   // cell list, near, far, total of the last execute
    int n_cu = 0;
};
