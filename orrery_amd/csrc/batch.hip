// Batches: many independent systems of uneven size stepped together.
//
// The reference steps one Simulation at a time (sim/simulation.cpp:26-64, one
// velocity_verlet_step per call, integrators/velocity_verlet.cpp:59-72), so a
// thousand small systems cost a thousand sequences of launches. Here a batch
// holds them all in HBM and one step of the whole batch is a fixed sequence of
// launches, whatever the number of systems:
//
//   state   ten FP64 arrays (x y z vx vy vz ax ay az m) over all particles,
//           the direct systems first, then the Barnes-Hut systems, each system
//           one contiguous range;
//   seg_of  system of each particle: direct systems are numbered 0..D-1 over
//           the direct range, Barnes-Hut systems 0..T-1 over the tree range;
//   tables  the direct tile and work-item tables, the tree segment tables,
//           per-system softening, G and kick/drift scales.
//
// A step is kick (all particles, each with its own system's scale), drift,
// then two branches on two streams: the direct systems (segmented Morton order,
// tile packing, one persistent work-item kernel, fixed-order reduction) and
// the Barnes-Hut systems (segmented Morton order, octree build, moments, walk),
// joined, then the closing kick. With use_graph the step is captured once as a
// hipGraph with the two branches in parallel, and replayed.
//
// A system's arithmetic depends only on its own particles: Morton order, cube
// and tree are per system, the direct source split depends on the system's size
// only, and every reduction runs in a fixed order. So a system gives the same
// bits alone in a batch as among any others.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <utility>
#include <vector>

#include "engine.hpp"

#pragma clang fp contract(off)

namespace orr {
namespace {

constexpr std::uint32_t kTile = 64;
constexpr std::uint32_t kChunk = 2048;  // particles per bounding-box chunk

// v += scale[system] a, the kick of integrators/velocity_verlet.cpp:20-38 (one
// multiply, one add, unfused) with each particle's own system's scale. Tree
// systems' scales follow the D direct ones.
__global__ __launch_bounds__(256) void batch_kick(std::uint64_t n, std::uint64_t n_direct,
                                                  std::uint32_t tree_base,
                                                  const std::uint32_t* __restrict__ seg_of,
                                                  const double* __restrict__ scale,
                                                  double* __restrict__ vx, double* __restrict__ vy,
                                                  double* __restrict__ vz,
                                                  const double* __restrict__ ax,
                                                  const double* __restrict__ ay,
                                                  const double* __restrict__ az) {
    const std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = scale[seg_of[i] + (i >= n_direct ? tree_base : 0u)];
    vx[i] = vx[i] + s * ax[i];
    vy[i] = vy[i] + s * ay[i];
    vz[i] = vz[i] + s * az[i];
}

template<typename T> T* carve(char*& p, std::size_t count) {
    T* out = reinterpret_cast<T*>(p);
    p += (count * sizeof(T) + 255) / 256 * 256;
    return out;
}

std::size_t padded(std::size_t bytes) { return (bytes + 255) / 256 * 256; }

}  // namespace
}  // namespace orr

struct orr_batch {
    orr_ctx* owner = nullptr;  // the caller's context: device
    orr_ctx* dctx = nullptr;   // direct branch: own stream and scratch
    orr_ctx* tctx = nullptr;   // Barnes-Hut branch: own stream and scratch
    orr_batch_config cfg{};
    std::vector<orr_batch_system> systems;  // caller's order
    std::vector<std::uint32_t> slot;        // caller's index -> internal
    std::vector<std::uint32_t> caller;      // internal -> caller's index
    std::vector<std::uint64_t> first;       // internal -> first particle (S + 1)
    std::uint32_t D = 0, T = 0;
    std::uint64_t n = 0, n_direct = 0;
    double tree_eps2 = 0.0, tree_G = 1.0;
    std::uint32_t tiles_per_item = 16, leaf_capacity = 16;
    bool tune_items = false, tune_leaf = false;

    orr::Buffer state, tables;
    // Device tables (carved from `tables`).
    std::uint32_t* seg_of = nullptr;
    double* scale = nullptr;  // [phase][half, full][S]
    uint4* dsys = nullptr;
    uint4* dtile = nullptr;
    uint4* ditem = nullptr;
    float* deps2 = nullptr;
    double* dG = nullptr;
    std::uint32_t *dstart = nullptr, *dchunks = nullptr, *dchunk_first = nullptr;
    std::uint32_t *tstart = nullptr, *tchunks = nullptr, *tchunk_first = nullptr;
    std::uint32_t dchunk_count = 0, tchunk_count = 0;
    std::uint32_t tiles = 0, items = 0, slots = 0;
    std::uint64_t direct_pairs = 0;
    bool safe = false;

    hipEvent_t fork = nullptr, join = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // The device buffers a captured step was recorded with (its launches hold
    // their addresses by value): the graph replays while these are unchanged.
    std::vector<std::pair<const void*, std::size_t>> graph_buffers;
    // dctx and tctx lent by the caller (orr::batch_create_on): kept, not freed.
    bool borrowed = false;

    double* array(int k) const { return state.as<double>() + static_cast<std::size_t>(k) * n; }
    std::uint32_t systems_internal() const { return D + T; }
};

namespace orr {
namespace {

void drop_graph(orr_batch* b) noexcept {
    if (b->exec != nullptr) (void)hipGraphExecDestroy(b->exec);
    if (b->graph != nullptr) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void sync_all(orr_batch* b) {
    b->tctx->sync();
    b->dctx->sync();
}

// Address and size of every device buffer a step touches (the batch's own and
// its two contexts' scratch), in address order: the ordering stage exchanges
// its two key and index buffers after an odd number of radix passes, which a
// captured step replays by itself, so only a new or resized buffer counts.
std::vector<std::pair<const void*, std::size_t>> step_buffers(const orr_batch* b) {
    std::vector<std::pair<const void*, std::size_t>> v;
    for (const Buffer* x : {&b->state, &b->tables}) v.emplace_back(x->ptr, x->bytes);
    for (orr_ctx* c : {b->dctx, b->tctx}) {
        for (const Buffer* x : c->buffers()) v.emplace_back(x->ptr, x->bytes);
    }
    std::sort(v.begin(), v.end());
    return v;
}

// Host tables for the current tiles_per_item; uploads everything and drops
// any captured step (its launches hold counts and pointers by value).
void plan(orr_batch* b) {
    drop_graph(b);
    const std::uint32_t S = b->systems_internal();
    std::vector<std::uint32_t> seg_of(b->n);
    std::vector<double> scale(6 * static_cast<std::size_t>(std::max<std::uint32_t>(S, 1)), 0.0);
    std::vector<uint4> dsys(b->D), dtile, ditem;
    std::vector<float> deps2(b->D);
    std::vector<double> dG(b->D);
    std::vector<std::uint32_t> dstart(b->D + 1), tstart(b->T + 1);
    std::vector<std::uint32_t> dchunks, tchunks, dchunk_first(b->D + 1), tchunk_first(b->T + 1);

    b->direct_pairs = 0;
    b->safe = false;
    std::uint32_t slots = 0;
    for (std::uint32_t s = 0; s < S; ++s) {
        const orr_batch_system& sys = b->systems[b->caller[s]];
        const bool direct = s < b->D;
        const std::uint64_t begin = b->first[s], end = b->first[s + 1];
        const std::uint32_t local = direct ? s : s - b->D;
        const std::uint64_t base = direct ? 0 : b->n_direct;
        for (std::uint64_t i = begin; i < end; ++i) seg_of[i] = local;
        // Scales exactly as the single-system stepper forms them: VV substeps of
        // dt (or w dt for Yoshida-4), kick dt/2 and drift dt.
        const double dt = sys.timestep;
        const double sub[3] = {b->cfg.integrator == 1 ? kOuterWeight * dt : dt, kInnerWeight * dt,
                               kOuterWeight * dt};
        for (int p = 0; p < 3; ++p) {
            scale[(2 * p) * S + s] = sub[p] / 2;
            scale[(2 * p + 1) * S + s] = sub[p];
        }
        auto& chunks = direct ? dchunks : tchunks;
        auto& chunk_first = direct ? dchunk_first : tchunk_first;
        chunk_first[local] = static_cast<std::uint32_t>(chunks.size() / 3);
        for (std::uint64_t c = begin; c < end; c += kChunk) {
            chunks.push_back(local);
            chunks.push_back(static_cast<std::uint32_t>(c - base));
            chunks.push_back(static_cast<std::uint32_t>(std::min(end, c + kChunk) - base));
        }
        chunk_first[local + 1] = static_cast<std::uint32_t>(chunks.size() / 3);
        if (!direct) {
            tstart[local] = static_cast<std::uint32_t>(begin - base);
            tstart[local + 1] = static_cast<std::uint32_t>(end - base);
            continue;
        }
        dstart[s] = static_cast<std::uint32_t>(begin);
        dstart[s + 1] = static_cast<std::uint32_t>(end);
        const std::uint64_t count = end - begin;
        b->direct_pairs += count * (count > 0 ? count - 1 : 0);
        b->safe = b->safe || (count > 0 && !(sys.softening_squared > 0));
        deps2[s] = static_cast<float>(sys.softening_squared);
        dG[s] = sys.gravitational_constant;
        const std::uint32_t ntiles = static_cast<std::uint32_t>((count + kTile - 1) / kTile);
        const std::uint32_t tile_base = static_cast<std::uint32_t>(dtile.size());
        const std::uint32_t splits = ntiles == 0 ? 0 : (ntiles + b->tiles_per_item - 1) / b->tiles_per_item;
        const std::uint32_t per = splits == 0 ? 0 : (ntiles + splits - 1) / splits;
        dsys[s] = make_uint4(static_cast<std::uint32_t>(begin), static_cast<std::uint32_t>(count),
                             tile_base, splits);
        for (std::uint32_t t = 0; t < ntiles; ++t) {
            const std::uint64_t tfirst = begin + static_cast<std::uint64_t>(t) * kTile;
            const std::uint32_t valid = static_cast<std::uint32_t>(std::min<std::uint64_t>(kTile, end - tfirst));
            dtile.push_back(make_uint4(static_cast<std::uint32_t>(tfirst), valid, s, slots));
            for (std::uint32_t k = 0; k < splits; ++k) {
                const std::uint32_t j0 = tile_base + k * per;
                const std::uint32_t j1 = std::min(tile_base + ntiles, j0 + per);
                ditem.push_back(make_uint4(tile_base + t, j0, j1, slots + k));
            }
            slots += splits;
        }
    }
    // Longest items first: with the strided schedule every wave starts on a long one.
    std::stable_sort(ditem.begin(), ditem.end(),
                     [](const uint4& a, const uint4& c) { return (a.z - a.y) > (c.z - c.y); });
    b->tiles = static_cast<std::uint32_t>(dtile.size());
    b->items = static_cast<std::uint32_t>(ditem.size());
    b->slots = slots;
    b->dchunk_count = static_cast<std::uint32_t>(dchunks.size() / 3);
    b->tchunk_count = static_cast<std::uint32_t>(tchunks.size() / 3);

    const std::size_t bytes =
        padded(4 * seg_of.size()) + padded(8 * scale.size()) + padded(16 * dsys.size()) +
        padded(16 * dtile.size()) + padded(16 * ditem.size()) + padded(4 * deps2.size()) +
        padded(8 * dG.size()) + padded(4 * dstart.size()) + padded(4 * dchunks.size()) +
        padded(4 * dchunk_first.size()) + padded(4 * tstart.size()) + padded(4 * tchunks.size()) +
        padded(4 * tchunk_first.size()) + 256;
    b->tables.reserve(bytes);
    char* p = static_cast<char*>(b->tables.ptr);
    b->seg_of = carve<std::uint32_t>(p, seg_of.size());
    b->scale = carve<double>(p, scale.size());
    b->dsys = carve<uint4>(p, dsys.size());
    b->dtile = carve<uint4>(p, dtile.size());
    b->ditem = carve<uint4>(p, ditem.size());
    b->deps2 = carve<float>(p, deps2.size());
    b->dG = carve<double>(p, dG.size());
    b->dstart = carve<std::uint32_t>(p, dstart.size());
    b->dchunks = carve<std::uint32_t>(p, dchunks.size());
    b->dchunk_first = carve<std::uint32_t>(p, dchunk_first.size());
    b->tstart = carve<std::uint32_t>(p, tstart.size());
    b->tchunks = carve<std::uint32_t>(p, tchunks.size());
    b->tchunk_first = carve<std::uint32_t>(p, tchunk_first.size());
    hipStream_t st = b->dctx->stream;
    auto put = [&](void* dst, const auto& v) {
        using T = typename std::decay_t<decltype(v)>::value_type;
        if (!v.empty()) ORR_HIP(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    };
    put(b->seg_of, seg_of);
    put(b->scale, scale);
    put(b->dsys, dsys);
    put(b->dtile, dtile);
    put(b->ditem, ditem);
    put(b->deps2, deps2);
    put(b->dG, dG);
    put(b->dstart, dstart);
    put(b->dchunks, dchunks);
    put(b->dchunk_first, dchunk_first);
    put(b->tstart, tstart);
    put(b->tchunks, tchunks);
    put(b->tchunk_first, tchunk_first);
    b->dctx->sync();
}

void kick_all(orr_batch* b, const double* scale, double* v0, double* v1, double* v2,
              const double* a0, const double* a1, const double* a2) {
    if (b->n == 0) return;
    hipLaunchKernelGGL(batch_kick, dim3(static_cast<std::uint32_t>((b->n + 255) / 256)), dim3(256),
                       0, b->dctx->stream, b->n, b->n_direct, b->D, b->seg_of, scale, v0, v1, v2,
                       a0, a1, a2);
    ORR_LAUNCHED("batch_kick");
}

void evaluate_direct(orr_batch* b) {
    if (b->n_direct == 0) return;
    orr_ctx* ctx = b->dctx;
    const SegmentView segs{b->D, b->dstart, b->seg_of, b->dchunks, b->dchunk_first, b->dchunk_count};
    morton_enqueue_segments(ctx, b->array(0), b->array(1), b->array(2), b->n_direct, segs);
    DirectBatchView v;
    v.count = b->n_direct;
    v.systems = b->D;
    v.tiles = b->tiles;
    v.items = b->items;
    v.slots = b->slots;
    v.sys = b->dsys;
    v.tile = b->dtile;
    v.item = b->ditem;
    v.seg_of = b->seg_of;
    v.eps2 = b->deps2;
    v.G = b->dG;
    v.safe = b->safe;
    direct_batch_enqueue(ctx, b->array(0), b->array(1), b->array(2), b->array(9),
                         ctx->order_index.as<std::uint32_t>(), v, b->array(6), b->array(7),
                         b->array(8));
}

void evaluate_tree(orr_batch* b) {
    const std::uint64_t base = b->n_direct;
    if (b->n == base) return;
    orr_tree_args a{};
    a.x = b->array(0) + base;
    a.y = b->array(1) + base;
    a.z = b->array(2) + base;
    a.m = b->array(9) + base;
    a.ax = b->array(6) + base;
    a.ay = b->array(7) + base;
    a.az = b->array(8) + base;
    a.count = b->n - base;
    a.softening_squared = b->tree_eps2;
    a.gravitational_constant = b->tree_G;
    a.opening_angle = b->cfg.opening_angle;
    a.leaf_capacity = b->leaf_capacity;
    a.quadrupole = b->cfg.quadrupole;
    const SegmentView segs{b->T, b->tstart, b->seg_of + base, b->tchunks, b->tchunk_first,
                           b->tchunk_count};
    tree_evaluate_segments(b->tctx, a, segs, nullptr, nullptr);
}

// Both branches, forked from and joined back into the direct stream.
void evaluate(orr_batch* b) {
    ORR_HIP(hipEventRecord(b->fork, b->dctx->stream));
    ORR_HIP(hipStreamWaitEvent(b->tctx->stream, b->fork, 0));
    evaluate_direct(b);
    evaluate_tree(b);
    ORR_HIP(hipEventRecord(b->join, b->tctx->stream));
    ORR_HIP(hipStreamWaitEvent(b->dctx->stream, b->join, 0));
}

// One step: velocity Verlet substeps (velocity_verlet.cpp:59-72), one for
// velocity Verlet, three weighted ones for Yoshida-4 (yoshida4.cpp:43-59).
void step_once(orr_batch* b) {
    const std::uint32_t S = b->systems_internal();
    const int substeps = b->cfg.integrator == 1 ? 3 : 1;
    for (int p = 0; p < substeps; ++p) {
        const double* half = b->scale + static_cast<std::size_t>(2 * p) * S;
        const double* full = b->scale + static_cast<std::size_t>(2 * p + 1) * S;
        kick_all(b, half, b->array(3), b->array(4), b->array(5), b->array(6), b->array(7), b->array(8));
        kick_all(b, full, b->array(0), b->array(1), b->array(2), b->array(3), b->array(4), b->array(5));
        evaluate(b);
        kick_all(b, half, b->array(3), b->array(4), b->array(5), b->array(6), b->array(7), b->array(8));
    }
}

void capture(orr_batch* b) {
    drop_graph(b);
    const auto before = step_buffers(b);
    hipStream_t st = b->dctx->stream;
    ORR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        step_once(b);
    } catch (...) {
        // A failure between the fork and the join leaves the Barnes-Hut
        // stream inside the capture: join it back first, or ending the
        // capture on the direct stream fails as unjoined and leaves that
        // stream capturing for good.
        hipStreamCaptureStatus status = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(b->tctx->stream, &status) == hipSuccess &&
            status != hipStreamCaptureStatusNone) {
            (void)hipEventRecord(b->join, b->tctx->stream);
            (void)hipStreamWaitEvent(st, b->join, 0);
        }
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g != nullptr) (void)hipGraphDestroy(g);
        throw;
    }
    ORR_HIP(hipStreamEndCapture(st, &b->graph));
    ORR_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    if (step_buffers(b) != before) {
        drop_graph(b);
        fail(ORR_ERR_RUNTIME, "batch: device buffers grew while a step was captured");
    }
    b->graph_buffers = before;
}

// Enqueues `steps` steps on the batch's streams. Only the first step after a
// change of plan or buffers waits (it runs eagerly, then the step is captured).
void step_enqueue(orr_batch* b, std::uint64_t steps) {
    if (b->n == 0) return;
    for (std::uint64_t k = 0; k < steps; ++k) {
        if (b->cfg.use_graph == 0 || !b->dctx->graphs) {
            step_once(b);
            continue;
        }
        if (b->exec != nullptr && step_buffers(b) == b->graph_buffers) {
            ORR_HIP(hipGraphLaunch(b->exec, b->dctx->stream));
            continue;
        }
        // An eager step first: it sizes every scratch buffer, so the capture
        // that follows allocates nothing.
        step_once(b);
        sync_all(b);
        capture(b);
    }
}

void step(orr_batch* b, std::uint64_t steps) {
    step_enqueue(b, steps);
    sync_all(b);
}

double time_ms(orr_batch* b, void (*body)(orr_batch*), int reps) {
    body(b);
    sync_all(b);
    const auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; ++r) body(b);
    sync_all(b);
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double, std::milli>(t1 - t0).count() / reps;
}

void autotune(orr_batch* b) {
    if (b->tune_items && b->n_direct > 0) {
        double best = 0.0;
        std::uint32_t pick = b->tiles_per_item;
        for (std::uint32_t cand : {4u, 8u, 16u, 32u, 64u}) {
            b->tiles_per_item = cand;
            plan(b);
            const double ms = time_ms(b, evaluate_direct, 5);
            if (pick == 0 || best == 0.0 || ms < best) {
                best = ms;
                pick = cand;
            }
        }
        b->tiles_per_item = pick;
        plan(b);
    }
    if (b->tune_leaf && b->n > b->n_direct) {
        double best = 0.0;
        std::uint32_t pick = b->leaf_capacity;
        for (std::uint32_t cand : {8u, 12u, 16u, 24u, 32u}) {
            b->leaf_capacity = cand;
            const double ms = time_ms(b, evaluate_tree, 3);
            if (best == 0.0 || ms < best) {
                best = ms;
                pick = cand;
            }
        }
        b->leaf_capacity = pick;
    }
    drop_graph(b);
    // Leave accelerations consistent with the chosen sizes.
    evaluate(b);
    sync_all(b);
}

std::uint32_t internal_of(orr_batch* b, std::uint32_t system) {
    require(system < b->systems.size(), "system index out of range");
    return b->slot[system];
}

void use_device(orr_batch* b) { ORR_HIP(hipSetDevice(b->owner->device)); }


void destroy(orr_batch* b) noexcept {
    if (b == nullptr) return;
    (void)hipSetDevice(b->owner->device);
    if (b->dctx != nullptr) (void)hipStreamSynchronize(b->dctx->stream);
    if (b->tctx != nullptr) (void)hipStreamSynchronize(b->tctx->stream);
    drop_graph(b);
    if (b->fork != nullptr) (void)hipEventDestroy(b->fork);
    if (b->join != nullptr) (void)hipEventDestroy(b->join);
    b->state.release();
    b->tables.release();
    if (!b->borrowed) {
        free_context(b->dctx);
        free_context(b->tctx);
    }
    delete b;
}

}  // namespace

std::uint64_t system_bytes_bound(std::uint64_t n, int solver, const orr_batch_config& cfg) {
    if (n == 0) return 0;
    const std::uint32_t below = cfg.direct_below == 0 ? 4096u : cfg.direct_below;
    const bool tree = solver == 1 || (solver == -1 && n >= below);
    // FP64 state, its table rows, and the diagnostics scratch.
    std::uint64_t total = grown_bytes(80 * n) + 4 * n + 256 + grown_bytes(diagnostics_scratch_bytes(n));
    // Ordering buffers, whichever way the system's ordering is laid out (alone,
    // or one of several segments with 2048-particle bounding-box chunks).
    const std::uint32_t chunks = static_cast<std::uint32_t>((n + kChunk - 1) / kChunk);
    const OrderSizes a = order_sizes(n, 2, chunks, tree), b = order_sizes(n, 1, 0, tree);
    total += grown_bytes(std::max(a.aos, b.aos)) + grown_bytes(std::max(a.bbox, b.bbox)) +
             2 * grown_bytes(a.codes) + 2 * grown_bytes(a.index) + grown_bytes(std::max(a.hist, b.hist));
    if (tree) {
        // The smallest leaf capacity the autotuner tries holds the most nodes.
        const std::uint64_t cap = cfg.leaf_capacity == 0 ? 8 : cfg.leaf_capacity;
        total += grown_bytes(tree_bytes(n, cap, cfg.quadrupole != 0, 1));
    } else {
        // The shortest work item the autotuner tries makes the most partial slots.
        const std::uint64_t tiles = (n + 63) / 64;
        const std::uint64_t per = cfg.tiles_per_item == 0 ? 4 : cfg.tiles_per_item;
        const std::uint64_t slots = tiles * ((tiles + per - 1) / per);
        total += grown_bytes(direct_batch_bytes(tiles, slots)) + 16 * (tiles + slots) + 12 * chunks;
    }
    return total;
}

void batch_step_enqueue(orr_batch* b, std::uint64_t steps) { step_enqueue(b, steps); }
void batch_wait(orr_batch* b) { sync_all(b); }
hipStream_t batch_stream(orr_batch* b) { return b->dctx->stream; }

void batch_refresh_enqueue(orr_batch* b) { evaluate(b); }

void batch_diagnostics_enqueue(orr_batch* b, std::uint32_t system, double* out) {
    const std::uint32_t s = internal_of(b, system);
    const orr_batch_system& sys = b->systems[system];
    const std::uint64_t begin = b->first[s], count = b->first[s + 1] - begin;
    diagnostics_enqueue(b->dctx, count, b->array(0) + begin, b->array(1) + begin,
                        b->array(2) + begin, b->array(3) + begin, b->array(4) + begin,
                        b->array(5) + begin, b->array(9) + begin, sys.softening_squared,
                        sys.gravitational_constant, out);
}

void batch_download_enqueue(orr_batch* b, std::uint32_t system, double* host) {
    const std::uint32_t s = internal_of(b, system);
    const std::uint64_t begin = b->first[s], count = b->first[s + 1] - begin;
    for (int k = 0; k < 6 && count > 0; ++k) {
        ORR_HIP(hipMemcpyAsync(host + static_cast<std::size_t>(k) * count, b->array(k) + begin,
                               count * sizeof(double), hipMemcpyDeviceToHost, b->dctx->stream));
    }
}

}  // namespace orr

using orr::guarded;
using orr::require;

namespace orr {

// orr_batch_create's body; with dctx and tctx given, the batch runs on those
// contexts (streams and scratch) and leaves them to the caller on destroy.
void batch_create(orr_ctx* ctx, const orr_batch_system* systems, uint32_t count,
                  const orr_batch_config* config, orr_ctx* lend_d, orr_ctx* lend_t,
                  orr_batch** out) {
    {
        require(ctx != nullptr && out != nullptr && config != nullptr, "null argument");
        require(count == 0 || systems != nullptr, "systems is null");
        *out = nullptr;
        require(config->integrator == 0 || config->integrator == 1,
                "integrator must be 0 (VV) or 1 (Yoshida-4)");
        ORR_HIP(hipSetDevice(ctx->device));
        std::unique_ptr<orr_batch, void (*)(orr_batch*)> b(new orr_batch, orr::destroy);
        b->owner = ctx;
        b->cfg = *config;
        b->systems.assign(systems, systems + count);
        const std::uint32_t below = config->direct_below == 0 ? 4096u : config->direct_below;
        std::vector<std::uint32_t> direct, tree;
        std::uint64_t total = 0;
        bool tree_params = false;
        for (std::uint32_t u = 0; u < count; ++u) {
            const orr_batch_system& s = systems[u];
            require(s.solver >= -1 && s.solver <= 1, "solver must be -1, 0 or 1");
            require(s.softening_squared >= 0, "softening squared must be >= 0");
            total += s.count;
            require(total < (std::uint64_t{1} << 31), "a batch holds fewer than 2^31 particles");
            const bool use_tree = s.count > 0 && (s.solver == 1 || (s.solver == -1 && s.count >= below));
            if (!use_tree) {
                direct.push_back(u);
                continue;
            }
            if (tree_params) {
                require(s.softening_squared == b->tree_eps2 && s.gravitational_constant == b->tree_G,
                        "the Barnes-Hut systems of a batch share softening and G");
            }
            tree_params = true;
            b->tree_eps2 = s.softening_squared;
            b->tree_G = s.gravitational_constant;
            tree.push_back(u);
        }
        b->D = static_cast<std::uint32_t>(direct.size());
        b->T = static_cast<std::uint32_t>(tree.size());
        b->slot.assign(count, 0);
        b->first.assign(1, 0);
        for (const auto* list : {&direct, &tree}) {
            for (std::uint32_t u : *list) {
                b->slot[u] = static_cast<std::uint32_t>(b->caller.size());
                b->caller.push_back(u);
                b->first.push_back(b->first.back() + systems[u].count);
            }
        }
        b->n = total;
        b->n_direct = b->first[b->D];
        b->tune_items = config->tiles_per_item == 0;
        b->tune_leaf = config->leaf_capacity == 0;
        b->tiles_per_item = b->tune_items ? 16u : config->tiles_per_item;
        b->leaf_capacity = b->tune_leaf ? 16u : config->leaf_capacity;
        if (lend_d != nullptr && lend_t != nullptr) {
            b->dctx = lend_d;
            b->tctx = lend_t;
            b->borrowed = true;
        } else {
            b->dctx = orr::make_context(ctx->device);
            b->tctx = orr::make_context(ctx->device);
        }
        // The owner's kernel settings carry over to the batch's private contexts.
        b->tctx->walk_targets_per_lane = ctx->walk_targets_per_lane;
        b->tctx->walk_lanes = ctx->walk_lanes;
        b->dctx->direct_refine = ctx->direct_refine;
        ORR_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
        ORR_HIP(hipEventCreateWithFlags(&b->join, hipEventDisableTiming));
        b->state.reserve(sizeof(double) * 10 * std::max<std::uint64_t>(total, 1));
        std::uint64_t largest = 1;
        for (std::uint32_t u = 0; u < count; ++u) largest = std::max<std::uint64_t>(largest, systems[u].count);
        b->dctx->diag_scratch.reserve(orr::diagnostics_scratch_bytes(largest));
        ORR_HIP(hipMemsetAsync(b->state.ptr, 0, b->state.bytes, b->dctx->stream));
        orr::plan(b.get());
        *out = b.release();
    }
}

orr_batch* batch_create_on(orr_ctx* owner, const orr_batch_system* systems, std::uint32_t count,
                           const orr_batch_config* config, orr_ctx* dctx, orr_ctx* tctx) {
    require(dctx != nullptr && tctx != nullptr, "null context");
    orr_batch* b = nullptr;
    batch_create(owner, systems, count, config, dctx, tctx, &b);
    return b;
}

}  // namespace orr

extern "C" {

int orr_batch_create(orr_ctx* ctx, const orr_batch_system* systems, uint32_t count,
                     const orr_batch_config* config, orr_batch** out) {
    return guarded([&] { orr::batch_create(ctx, systems, count, config, nullptr, nullptr, out); });
}

int orr_batch_destroy(orr_batch* batch) {
    return guarded([&] { orr::destroy(batch); });
}

int orr_batch_info_get(orr_batch* b, orr_batch_info* out) {
    return guarded([&] {
        require(b != nullptr && out != nullptr, "null argument");
        *out = orr_batch_info{};
        out->systems = static_cast<uint32_t>(b->systems.size());
        out->direct_systems = b->D;
        out->tree_systems = b->T;
        out->tiles_per_item = b->tiles_per_item;
        out->leaf_capacity = b->leaf_capacity;
        out->direct_tiles = b->tiles;
        out->direct_items = b->items;
        out->graph = b->exec != nullptr ? 1 : 0;
        out->particles = b->n;
        out->direct_particles = b->n_direct;
        out->tree_particles = b->n - b->n_direct;
        out->direct_pairs = b->direct_pairs;
        out->device_bytes = b->state.bytes + b->tables.bytes + orr::ctx_device_bytes(b->dctx) +
                            orr::ctx_device_bytes(b->tctx);
    });
}

int orr_batch_bytes_bound(const orr_batch_system* systems, uint32_t count,
                          const orr_batch_config* config, uint64_t* bytes_out) {
    return guarded([&] {
        require(config != nullptr && bytes_out != nullptr, "null argument");
        require(count == 0 || systems != nullptr, "systems is null");
        std::uint64_t total = 16 * 256;  // table padding
        for (std::uint32_t u = 0; u < count; ++u) {
            require(systems[u].solver >= -1 && systems[u].solver <= 1, "solver must be -1, 0 or 1");
            total += orr::system_bytes_bound(systems[u].count, systems[u].solver, *config) + 256;
        }
        *bytes_out = total;
    });
}

int orr_batch_upload(orr_batch* b, uint32_t system, const double* x, const double* y,
                     const double* z, const double* vx, const double* vy, const double* vz,
                     const double* m) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        const std::uint32_t s = orr::internal_of(b, system);
        orr::use_device(b);
        const std::uint64_t begin = b->first[s], count = b->first[s + 1] - begin;
        const double* src[10] = {x, y, z, vx, vy, vz, nullptr, nullptr, nullptr, m};
        for (int k = 0; k < 10; ++k) {
            if (src[k] == nullptr || count == 0) continue;
            ORR_HIP(hipMemcpyAsync(b->array(k) + begin, src[k], count * sizeof(double),
                                   hipMemcpyHostToDevice, b->dctx->stream));
        }
        orr::sync_all(b);
    });
}

int orr_batch_download(orr_batch* b, uint32_t system, double* x, double* y, double* z, double* vx,
                       double* vy, double* vz, double* ax, double* ay, double* az) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        const std::uint32_t s = orr::internal_of(b, system);
        orr::use_device(b);
        orr::sync_all(b);
        const std::uint64_t begin = b->first[s], count = b->first[s + 1] - begin;
        double* dst[9] = {x, y, z, vx, vy, vz, ax, ay, az};
        for (int k = 0; k < 9; ++k) {
            if (dst[k] == nullptr || count == 0) continue;
            ORR_HIP(hipMemcpyAsync(dst[k], b->array(k) + begin, count * sizeof(double),
                                   hipMemcpyDeviceToHost, b->dctx->stream));
        }
        b->dctx->sync();
    });
}

int orr_batch_copy_system(orr_batch* dst, uint32_t dst_system, orr_batch* src, uint32_t src_system) {
    return guarded([&] {
        require(dst != nullptr && src != nullptr, "null batch");
        const std::uint32_t d = orr::internal_of(dst, dst_system);
        const std::uint32_t s = orr::internal_of(src, src_system);
        const std::uint64_t count = dst->first[d + 1] - dst->first[d];
        require(count == src->first[s + 1] - src->first[s], "systems differ in size");
        require(dst->owner->device == src->owner->device, "batches on different devices");
        orr::use_device(dst);
        orr::sync_all(src);
        orr::sync_all(dst);
        if (count == 0) return;
        for (int k = 0; k < 10; ++k) {
            ORR_HIP(hipMemcpyAsync(dst->array(k) + dst->first[d], src->array(k) + src->first[s],
                                   count * sizeof(double), hipMemcpyDeviceToDevice,
                                   dst->dctx->stream));
        }
        dst->dctx->sync();
    });
}

int orr_batch_refresh(orr_batch* b) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        orr::use_device(b);
        orr::evaluate(b);
        orr::sync_all(b);
    });
}

int orr_batch_step(orr_batch* b, uint64_t steps) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        orr::use_device(b);
        orr::step(b, steps);
    });
}

int orr_batch_autotune(orr_batch* b) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        orr::use_device(b);
        orr::autotune(b);
    });
}

int orr_batch_diagnostics(orr_batch* b, uint32_t system, double out[8]) {
    return guarded([&] {
        require(b != nullptr && out != nullptr, "null argument");
        const std::uint32_t s = orr::internal_of(b, system);
        orr::use_device(b);
        orr::sync_all(b);
        (void)s;
        orr::batch_diagnostics_enqueue(b, system, out);
        orr::sync_all(b);
    });
}

int orr_batch_counts(orr_batch* b, uint32_t system, uint64_t out[2]) {
    return guarded([&] {
        require(b != nullptr && out != nullptr, "null argument");
        const std::uint32_t s = orr::internal_of(b, system);
        orr::use_device(b);
        orr::sync_all(b);
        const std::uint64_t count = b->first[s + 1] - b->first[s];
        if (s < b->D) {
            out[0] = count * (count > 0 ? count - 1 : 0);
            out[1] = 0;
            return;
        }
        std::vector<std::uint64_t> c(4 * static_cast<std::size_t>(b->T));
        orr::tree_copy_segment_counts(b->tctx, c.data());
        out[0] = c[4 * (s - b->D)];
        out[1] = c[4 * (s - b->D) + 1];
    });
}

}  // extern "C"
