// Context, device memory, errors and phase timing of the engine.
//
// Replaces the reference's CUDA runtime plumbing (backend/cuda_device.cpp:71-115,
// backend/cuda_memory.hpp:77-348): device discovery, RAII device arrays and the
// status-to-error mapping, written against the HIP runtime directly.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdlib>
#include <cstring>
#include <memory>
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

std::atomic<std::uint64_t> g_reallocations{0};

std::size_t grown_bytes(std::size_t wanted) { return wanted == 0 ? 0 : wanted + wanted / 4 + 256; }

void Buffer::reserve(std::size_t wanted) {
    if (wanted <= bytes) return;
    g_reallocations.fetch_add(1, std::memory_order_relaxed);
    release();
    // Grow by a quarter beyond the request so that slowly growing systems do
    // not reallocate on every step.
    const std::size_t grown = grown_bytes(wanted);
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

std::size_t ctx_device_bytes(const orr_ctx* ctx) {
    std::size_t total = 0;
    for (const Buffer* b : const_cast<orr_ctx*>(ctx)->buffers()) total += b->bytes;
    return total;
}

void drop_captured(orr_ctx::Captured& g) noexcept {
    if (g.exec != nullptr) (void)hipGraphExecDestroy(g.exec);
    if (g.graph != nullptr) (void)hipGraphDestroy(g.graph);
    g = orr_ctx::Captured{};
}

void run_captured(orr_ctx* ctx, orr_ctx::Captured& g, std::vector<unsigned char> key,
                  const std::function<void()>& enqueue, bool untimed) {
    if (ctx->capturing) {
        enqueue();
        return;
    }
    if (!ctx->graphs || (ctx->timing && !untimed)) {
        drop_captured(g);
        enqueue();
        return;
    }
    // Phase timers stay off inside (and are restored however this returns).
    struct Untimed {
        orr_ctx* ctx;
        bool was;
        ~Untimed() { ctx->timing = was; }
    } restore{ctx, ctx->timing};
    ctx->timing = false;
    // The sort form changes the launches, not the arguments: part of the key.
    key_add(key, ctx->sort_mode);
    if (!(g.seen && g.key == key && g.epoch == g_reallocations.load())) {
        drop_captured(g);
        enqueue();
        g.key = std::move(key);
        g.epoch = g_reallocations.load();
        g.seen = true;
        return;
    }
    if (g.exec == nullptr) {
        ORR_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        hipGraph_t graph = nullptr;
        ctx->capturing = true;
        try {
            enqueue();
        } catch (...) {
            ctx->capturing = false;
            (void)hipStreamEndCapture(ctx->stream, &graph);
            if (graph != nullptr) (void)hipGraphDestroy(graph);
            drop_captured(g);
            throw;
        }
        ctx->capturing = false;
        ORR_HIP(hipStreamEndCapture(ctx->stream, &graph));
        g.graph = graph;
        const hipError_t made = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        if (made != hipSuccess || g.epoch != g_reallocations.load()) {
            // Nothing may allocate while capturing; if something did, run eagerly.
            drop_captured(g);
            enqueue();
            return;
        }
    }
    ORR_HIP(hipGraphLaunch(g.exec, ctx->stream));
    ++ctx->totals.graph_launches;
}

orr_ctx* make_context(int device) {
    ORR_HIP(hipSetDevice(device));
    auto ctx = std::make_unique<orr_ctx>();
    ctx->device = device;
    hipDeviceProp_t p{};
    ORR_HIP(hipGetDeviceProperties(&p, device));
    ctx->compute_units = p.multiProcessorCount;
    // Default priority on purpose: streams of mixed priority are not used.
    ORR_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    const char* graphs = std::getenv("ORR_GRAPHS");
    ctx->graphs = !(graphs != nullptr && std::strcmp(graphs, "0") == 0);
    return ctx.release();
}

void ensure_aux(orr_ctx* ctx) {
    if (ctx->aux != nullptr) return;
    // Default priority, as the main stream.
    ORR_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
    ORR_HIP(hipEventCreateWithFlags(&ctx->aux_fork, hipEventDisableTiming));
    ORR_HIP(hipEventCreateWithFlags(&ctx->aux_join, hipEventDisableTiming));
}

void free_context(orr_ctx* ctx) noexcept {
    if (ctx == nullptr) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    drop_captured(ctx->tree_graph);
    drop_captured(ctx->direct_graph);
    for (Buffer* b : ctx->buffers()) b->release();
    for (auto& p : ctx->pending) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    for (hipEvent_t e : ctx->free_events) (void)hipEventDestroy(e);
    if (ctx->aux != nullptr) {
        (void)hipStreamSynchronize(ctx->aux);
        (void)hipStreamDestroy(ctx->aux);
    }
    if (ctx->aux_fork != nullptr) (void)hipEventDestroy(ctx->aux_fork);
    if (ctx->aux_join != nullptr) (void)hipEventDestroy(ctx->aux_join);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
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
