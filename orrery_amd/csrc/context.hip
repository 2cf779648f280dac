// Context, device memory, errors and phase timing of the engine.
//
// Replaces the reference's CUDA runtime plumbing (backend/cuda_device.cpp:71-115,
// backend/cuda_memory.hpp:77-348): device discovery, RAII device arrays and the
// status-to-error mapping, written against the HIP runtime directly.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "engine.hpp"

namespace orr {

void fail(int status, const std::string& what) { throw Error(status, what); }

void check_hip(hipError_t e, const char* what, const char* file, int line) {
    if (e == hipSuccess) return;
    std::string msg = std::string(what) + ": " + hipGetErrorString(e) + " (" + file + ":" +
                      std::to_string(line) + ")";
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) fail(ORR_ERR_OUT_OF_MEMORY, msg);
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) fail(ORR_ERR_NO_DEVICE, msg);
    fail(ORR_ERR_RUNTIME, msg);
}

void Buffer::reserve(std::size_t wanted) {
    if (wanted <= bytes) return;
    release();
    // Grow by a quarter beyond the request so that slowly growing systems do
    // not reallocate on every step.
    const std::size_t grown = wanted + wanted / 4 + 256;
    void* p = nullptr;
    ORR_HIP(hipMalloc(&p, grown));
    ptr = p;
    bytes = grown;
}

void Buffer::release() noexcept {
    if (ptr != nullptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

}  // namespace orr

hipEvent_t orr_ctx::take_event() {
    if (!free_events.empty()) {
        hipEvent_t e = free_events.back();
        free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    ORR_HIP(hipEventCreate(&e));
    return e;
}

void orr_ctx::begin(int, hipEvent_t& start) {
    if (!timing) return;
    start = take_event();
    ORR_HIP(hipEventRecord(start, stream));
}

void orr_ctx::end(int phase, hipEvent_t start) {
    if (!timing || start == nullptr) return;
    hipEvent_t stop = take_event();
    ORR_HIP(hipEventRecord(stop, stream));
    pending.push_back(Pending{phase, start, stop});
}

void orr_ctx::sync() {
    ORR_HIP(hipStreamSynchronize(stream));
    for (const Pending& p : pending) {
        float ms = 0.f;
        ORR_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
        totals.milliseconds[p.phase] += ms;
        totals.launches[p.phase] += 1;
        free_events.push_back(p.start);
        free_events.push_back(p.stop);
    }
    pending.clear();
}
