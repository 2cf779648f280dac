// extern "C" surface of the engine (include/orrery_hip.h).
//
// Each entry point validates its arguments, runs the C++ implementation and
// maps any orr::Error to its status with a thread-local message: the wasm C
// ABI's convention (wasm/orrery_wasm.h:86-94) applied to the device engine.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "engine.hpp"

namespace {

thread_local std::string g_last_error;

using orr::guarded;
using orr::require;

void use_device(orr_ctx* ctx) { ORR_HIP(hipSetDevice(ctx->device)); }

}  // namespace

void orr::set_last_error(const char* what) { g_last_error = what; }

struct orr_system {
    orr_ctx* ctx = nullptr;
    std::uint64_t count = 0;
    orr::Buffer state;   // x y z vx vy vz ax ay az m, FP64, count each
    orr::Buffer stages;  // RK4 only: x0 (3), v0 (3), position sums (3), velocity sums (3)
    orr_ctx::Captured step_graph;  // one whole step (kicks, drifts, field), replayed
    double* array(int k) const { return state.as<double>() + static_cast<std::size_t>(k) * count; }
    double* stage(int k) const { return stages.as<double>() + static_cast<std::size_t>(k) * count; }
};

std::uint64_t orr::system_count(const orr_system* s) { return s->count; }

void orr::system_download(orr_system* s, int first, int arrays, double* host) {
    ORR_HIP(hipSetDevice(s->ctx->device));
    if (s->count == 0) return;
    ORR_HIP(hipMemcpyAsync(host, s->array(first), sizeof(double) * arrays * s->count,
                           hipMemcpyDeviceToHost, s->ctx->stream));
    s->ctx->sync();
}

void orr::system_upload(orr_system* s, int first, int arrays, const double* host) {
    ORR_HIP(hipSetDevice(s->ctx->device));
    if (s->count == 0) return;
    ORR_HIP(hipMemcpyAsync(s->array(first), host, sizeof(double) * arrays * s->count,
                           hipMemcpyHostToDevice, s->ctx->stream));
    s->ctx->sync();
}

namespace {

// Evaluates a(x); with `kick`, also v += kick->s a when the solver can fuse it
// into its fold (direct), which the return value reports.
bool field_eval(orr_system* s, const orr_step_config& c, const orr::FusedKick* kick = nullptr) {
    if (c.solver == 1) {
        orr_tree_args t{};
        t.x = s->array(0);
        t.y = s->array(1);
        t.z = s->array(2);
        t.m = s->array(9);
        t.ax = s->array(6);
        t.ay = s->array(7);
        t.az = s->array(8);
        t.count = s->count;
        t.softening_squared = c.softening_squared;
        t.gravitational_constant = c.gravitational_constant;
        t.opening_angle = c.opening_angle;
        t.leaf_capacity = c.leaf_capacity;
        t.quadrupole = c.quadrupole;
        orr::tree_evaluate(s->ctx, t, nullptr);
        return false;
    }
    {
        orr_direct_args d{};
        d.x = s->array(0);
        d.y = s->array(1);
        d.z = s->array(2);
        d.m = s->array(9);
        d.ax = s->array(6);
        d.ay = s->array(7);
        d.az = s->array(8);
        d.count = s->count;
        d.softening_squared = c.softening_squared;
        d.gravitational_constant = c.gravitational_constant;
        d.spatial_order = 1;
        orr::direct_evaluate(s->ctx, d, kick != nullptr ? *kick : orr::FusedKick{});
        return kick != nullptr;
    }
}

// velocity_verlet.cpp:59-72: kick half, drift, evaluate, kick half.
void vv_substep(orr_system* s, const orr_step_config& c, double dt) {
    const double half = dt / 2;
    const std::uint64_t n = s->count;
    // The first kick and the drift as one launch; the second kick inside the
    // direct fold where it can be (the same bits as four separate launches).
    orr::kick_drift(s->ctx, n, s->array(0), s->array(1), s->array(2), s->array(3), s->array(4),
                    s->array(5), s->array(6), s->array(7), s->array(8), half, dt);
    const orr::FusedKick k{s->array(3), s->array(4), s->array(5), half};
    if (!field_eval(s, c, &k)) {
        orr::kick(s->ctx, n, s->array(3), s->array(4), s->array(5), s->array(6), s->array(7),
                  s->array(8), half);
    }
}

using orr::kInnerWeight;
using orr::kOuterWeight;

// runge_kutta4.cpp:64-139: three stages (dt/2 weight 2, dt/2 weight 2, dt weight
// 1) after the first, combined with dt/6; a fourth evaluation restores a(x).
void rk4_step(orr_system* s, const orr_step_config& c, double dt) {
    const std::uint64_t n = s->count;
    s->stages.reserve(sizeof(double) * 12 * n);
    orr::Rk4Arrays r{};
    for (int k = 0; k < 3; ++k) {
        r.x[k] = s->array(k);
        r.v[k] = s->array(3 + k);
        r.a[k] = s->array(6 + k);
        r.x0[k] = s->stage(k);
        r.v0[k] = s->stage(3 + k);
        r.sx[k] = s->stage(6 + k);
        r.sv[k] = s->stage(9 + k);
    }
    orr::rk4(s->ctx, n, r, 0, 0.0);
    const double half_step = dt / 2;
    const double scales[3] = {half_step, half_step, dt};
    const double weights[3] = {2, 2, 1};
    for (int k = 0; k < 3; ++k) {
        orr::rk4(s->ctx, n, r, 1, scales[k]);
        field_eval(s, c);
        orr::rk4(s->ctx, n, r, 2, weights[k]);
    }
    const double sixth_step = dt / 6;
    orr::rk4(s->ctx, n, r, 3, sixth_step);
    field_eval(s, c);
}

void check_config(const orr_step_config* c) {
    require(c != nullptr, "step config is null");
    require(c->integrator >= 0 && c->integrator <= 2,
            "integrator must be 0 (VV), 1 (Yoshida-4) or 2 (RK4)");
    require(c->solver == 0 || c->solver == 1, "solver must be 0 (direct) or 1 (Barnes-Hut)");
    require(c->softening_squared >= 0, "softening squared must be >= 0");
}

}  // namespace

extern "C" {

const char* orr_last_error(void) { return g_last_error.c_str(); }

int orr_device_count(int32_t* count) {
    return guarded([&] {
        require(count != nullptr, "count is null");
        int n = 0;
        const hipError_t e = hipGetDeviceCount(&n);
        if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) n = 0;
        else ORR_HIP(e);
        *count = n;
    });
}

int orr_device_describe(int32_t device, orr_device_desc* out) {
    return guarded([&] {
        require(out != nullptr, "out is null");
        hipDeviceProp_t p{};
        ORR_HIP(hipGetDeviceProperties(&p, device));
        std::memset(out, 0, sizeof(*out));
        std::strncpy(out->name, p.name, sizeof(out->name) - 1);
        std::strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
        out->compute_units = p.multiProcessorCount;
        out->wavefront_size = p.warpSize;
        out->max_threads_per_block = p.maxThreadsPerBlock;
        out->clock_khz = p.clockRate;
        out->global_memory_bytes = p.totalGlobalMem;
        out->lds_per_block_bytes = p.sharedMemPerBlock;
        out->l2_bytes = static_cast<uint64_t>(p.l2CacheSize);
    });
}

int orr_ctx_create(int32_t device, orr_ctx** out) {
    return guarded([&] {
        require(out != nullptr, "out is null");
        *out = nullptr;
        int n = 0;
        const hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n == 0) orr::fail(ORR_ERR_NO_DEVICE, "no HIP device");
        require(device >= 0 && device < n, "device index out of range");
        *out = orr::make_context(device);
    });
}

int orr_ctx_destroy(orr_ctx* ctx) {
    return guarded([&] { orr::free_context(ctx); });
}

int orr_malloc(orr_ctx* ctx, uint64_t bytes, void** out) {
    return guarded([&] {
        require(ctx != nullptr && out != nullptr, "null argument");
        use_device(ctx);
        *out = nullptr;
        ORR_HIP(hipMalloc(out, bytes == 0 ? 1 : bytes));
    });
}

int orr_free(orr_ctx* ctx, void* p) {
    return guarded([&] {
        require(ctx != nullptr, "null context");
        use_device(ctx);
        if (p != nullptr) ORR_HIP(hipFree(p));
    });
}

int orr_upload(orr_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    return guarded([&] {
        require(ctx != nullptr && (bytes == 0 || (dst != nullptr && src != nullptr)), "null argument");
        use_device(ctx);
        if (bytes == 0) return;
        ORR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        ctx->sync();
    });
}

int orr_download(orr_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    return guarded([&] {
        require(ctx != nullptr && (bytes == 0 || (dst != nullptr && src != nullptr)), "null argument");
        use_device(ctx);
        if (bytes == 0) return;
        ORR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
    });
}

int orr_memset(orr_ctx* ctx, void* dst, int32_t value, uint64_t bytes) {
    return guarded([&] {
        require(ctx != nullptr && (bytes == 0 || dst != nullptr), "null argument");
        use_device(ctx);
        if (bytes == 0) return;
        ORR_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
        ctx->sync();
    });
}

int orr_direct_eval(orr_ctx* ctx, const orr_direct_args* a) {
    return guarded([&] {
        require(ctx != nullptr && a != nullptr, "null argument");
        require(a->softening_squared >= 0, "softening squared must be >= 0");
        if (a->count == 0) return;
        require(a->x && a->y && a->z && a->m && a->ax && a->ay && a->az, "null particle array");
        use_device(ctx);
        orr::direct_evaluate(ctx, *a);
        ctx->sync();
    });
}

int orr_direct_eval_static(orr_ctx* ctx, const orr_direct_args* a) {
    return guarded([&] {
        require(ctx != nullptr && a != nullptr, "null argument");
        require(a->softening_squared >= 0, "softening squared must be >= 0");
        if (a->count == 0) return;
        require(a->x && a->y && a->z && a->m && a->ax && a->ay && a->az, "null particle array");
        use_device(ctx);
        orr::direct_evaluate_static(ctx, *a);
        // A plain stream synchronise: ctx->sync() would also fold pending
        // timing events into the context's host tables.
        ORR_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int orr_direct_static_limits(uint64_t* max_count, uint64_t* scratch_bytes) {
    return guarded([&] {
        require(max_count != nullptr && scratch_bytes != nullptr, "null argument");
        *max_count = orr::direct_static_max_count();
        *scratch_bytes = orr::direct_static_scratch_bytes();
    });
}

int orr_ctx_set_direct_config(orr_ctx* ctx, const orr_direct_config* c) {
    return guarded([&] {
        require(ctx != nullptr && c != nullptr, "null argument");
        require(c->targets_per_lane == 0 || c->targets_per_lane == 1 || c->targets_per_lane == 2 ||
                    c->targets_per_lane == 4,
                "targets_per_lane must be 1, 2 or 4");
        require(c->refine == 0 || c->refine == 1 || c->refine == -1, "refine must be 0 or 1 (-1 keeps)");
        require(c->pair_mode >= 0 && c->pair_mode <= 3, "pair_mode must be 1, 2 or 3 (0 keeps)");
        require(c->equal_masses >= 0 && c->equal_masses <= 2, "equal_masses must be 1 or 2 (0 keeps)");
        require(c->small_order >= 0 && c->small_order <= 2, "small_order must be 1 or 2 (0 keeps)");
        if (c->small_order != 0) ctx->direct_small_order = c->small_order;
        if (c->equal_masses != 0) ctx->direct_equal_masses = c->equal_masses;
        if (c->targets_per_lane != 0) ctx->direct_targets_per_lane = c->targets_per_lane;
        if (c->pair_mode != 0) ctx->direct_pair_mode = c->pair_mode;
        if (c->refine >= 0) ctx->direct_refine = c->refine;
        if (c->block_target != 0) ctx->direct_block_target = c->block_target;
    });
}

int orr_ctx_set_tree_config(orr_ctx* ctx, const orr_tree_config* config) {
    return guarded([&] {
        require(ctx != nullptr && config != nullptr, "null argument");
        const int r = config->targets_per_lane;
        require(r >= -1 && r <= 2, "targets_per_lane must be 1, 2, 0 (keep) or -1 (by size)");
        const int l = config->lanes;
        require(l == -1 || l == 0 || l == 32 || l == 64, "lanes must be 32, 64, 0 (keep) or -1 (by size)");
        if (r == -1) ctx->walk_targets_per_lane = 0;
        if (r > 0) ctx->walk_targets_per_lane = r;
        require(config->sort_mode >= 0 && config->sort_mode <= 2, "sort_mode must be 1, 2 or 0 (keep)");
        if (l == -1) ctx->walk_lanes = 0;
        if (l > 0) ctx->walk_lanes = l;
        if (config->sort_mode != 0) ctx->sort_mode = config->sort_mode;
    });
}

int orr_ctx_get_tree_config(orr_ctx* ctx, orr_tree_config* config) {
    return guarded([&] {
        require(ctx != nullptr && config != nullptr, "null argument");
        *config = orr_tree_config{};
        // -1 while the choice is by size, as orr_ctx_set_tree_config spells it.
        config->targets_per_lane = ctx->walk_targets_per_lane != 0 ? ctx->walk_targets_per_lane : -1;
        config->lanes = ctx->walk_lanes != 0 ? ctx->walk_lanes : -1;
        config->sort_mode = ctx->sort_mode;
    });
}

int orr_ctx_get_direct_config(orr_ctx* ctx, orr_direct_config* c) {
    return guarded([&] {
        require(ctx != nullptr && c != nullptr, "null argument");
        c->targets_per_lane = ctx->direct_targets_per_lane;
        c->refine = ctx->direct_refine;
        c->block_target = ctx->direct_block_target;
        c->pair_mode = ctx->direct_pair_mode;
        c->equal_masses = ctx->direct_equal_masses;
        c->small_order = ctx->direct_small_order;
    });
}

int orr_tree_eval(orr_ctx* ctx, const orr_tree_args* a, orr_tree_stats* stats) {
    return guarded([&] {
        require(ctx != nullptr && a != nullptr, "null argument");
        require(a->softening_squared >= 0, "softening squared must be >= 0");
        if (a->count > 0) {
            require(a->x && a->y && a->z && a->m && a->ax && a->ay && a->az, "null particle array");
        }
        use_device(ctx);
        orr::tree_evaluate(ctx, *a, stats);
        ctx->sync();
    });
}

int orr_morton_order(orr_ctx* ctx, const double* x, const double* y, const double* z,
                     uint64_t n, uint64_t* codes_out, uint32_t* index_out, double cube_out[4]) {
    return guarded([&] {
        require(ctx != nullptr, "null context");
        use_device(ctx);
        if (n == 0) {
            if (cube_out) cube_out[0] = cube_out[1] = cube_out[2] = cube_out[3] = 0.0;
            return;
        }
        require(x && y && z && codes_out && index_out, "null array");
        orr::morton_enqueue(ctx, x, y, z, n);
        ORR_HIP(hipMemcpyAsync(codes_out, ctx->order_codes.ptr, n * sizeof(uint64_t),
                               hipMemcpyDeviceToDevice, ctx->stream));
        ORR_HIP(hipMemcpyAsync(index_out, ctx->order_index.ptr, n * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, ctx->stream));
        double cube[4];
        orr::morton_cube(ctx, cube);
        if (cube_out) std::memcpy(cube_out, cube, sizeof(cube));
    });
}

int orr_tree_nodes(orr_ctx* ctx, double* mass, double* cx, double* cy, double* cz, double* acc,
                   uint64_t* next, uint64_t* first, uint64_t* count) {
    return guarded([&] {
        require(ctx != nullptr, "null context");
        use_device(ctx);
        orr::tree_copy_nodes(ctx, mass, cx, cy, cz, acc, next, first, count);
    });
}

int orr_tree_quadrupoles(orr_ctx* ctx, double* out) {
    return guarded([&] {
        require(ctx != nullptr && out != nullptr, "null argument");
        use_device(ctx);
        orr::tree_copy_quadrupoles(ctx, out);
    });
}

int orr_probe(orr_ctx* ctx, int32_t kind, uint64_t size, double* rate, double* ms) {
    return guarded([&] {
        require(ctx != nullptr && rate != nullptr && ms != nullptr, "null argument");
        require(kind >= 0 && kind <= 4, "probe kind must be 0..4");
        require(size > 0, "probe size must be > 0");
        const bool iterations = kind <= 1 || kind == 4;
        require(!iterations || size <= (uint64_t{1} << 24), "probe iterations must be <= 2^24");
        require(iterations || size <= (uint64_t{1} << 34), "probe buffer must be <= 16 GiB");
        use_device(ctx);
        ctx->sync();
        orr::probe(ctx, kind, size, rate, ms);
    });
}

int orr_ctx_set_timing(orr_ctx* ctx, int32_t enabled) {
    return guarded([&] {
        require(ctx != nullptr, "null context");
        use_device(ctx);
        ctx->sync();
        ctx->timing = enabled != 0;
    });
}

int orr_ctx_timings(orr_ctx* ctx, orr_timings* out) {
    return guarded([&] {
        require(ctx != nullptr && out != nullptr, "null argument");
        use_device(ctx);
        ctx->sync();
        *out = ctx->totals;
    });
}

int orr_ctx_reset_timings(orr_ctx* ctx) {
    return guarded([&] {
        require(ctx != nullptr, "null context");
        use_device(ctx);
        ctx->sync();
        ctx->totals = orr_timings{};
    });
}

int orr_system_create(orr_ctx* ctx, uint64_t count, orr_system** out) {
    return guarded([&] {
        require(ctx != nullptr && out != nullptr, "null argument");
        *out = nullptr;
        require(count <= (uint64_t{1} << 31), "more than 2^31 particles");
        use_device(ctx);
        auto s = std::make_unique<orr_system>();
        s->ctx = ctx;
        s->count = count;
        s->state.reserve(sizeof(double) * 10 * (count == 0 ? 1 : count));
        ORR_HIP(hipMemsetAsync(s->state.ptr, 0, s->state.bytes, ctx->stream));
        ctx->sync();
        *out = s.release();
    });
}

int orr_system_count(orr_system* s, uint64_t* count) {
    return guarded([&] {
        require(s != nullptr && count != nullptr, "null argument");
        *count = s->count;
    });
}

int orr_system_destroy(orr_system* s) {
    return guarded([&] {
        if (s == nullptr) return;
        use_device(s->ctx);
        s->ctx->sync();
        orr::drop_captured(s->step_graph);
        s->state.release();
        s->stages.release();
        delete s;
    });
}

int orr_system_upload(orr_system* s, const double* x, const double* y, const double* z,
                      const double* vx, const double* vy, const double* vz, const double* m) {
    return guarded([&] {
        require(s != nullptr, "null system");
        use_device(s->ctx);
        const double* src[10] = {x, y, z, vx, vy, vz, nullptr, nullptr, nullptr, m};
        for (int k = 0; k < 10; ++k) {
            if (src[k] == nullptr || s->count == 0) continue;
            ORR_HIP(hipMemcpyAsync(s->array(k), src[k], s->count * sizeof(double),
                                   hipMemcpyHostToDevice, s->ctx->stream));
        }
        s->ctx->sync();
    });
}

int orr_system_download(orr_system* s, double* x, double* y, double* z, double* vx, double* vy,
                        double* vz, double* ax, double* ay, double* az) {
    return guarded([&] {
        require(s != nullptr, "null system");
        use_device(s->ctx);
        double* dst[9] = {x, y, z, vx, vy, vz, ax, ay, az};
        for (int k = 0; k < 9; ++k) {
            if (dst[k] == nullptr || s->count == 0) continue;
            ORR_HIP(hipMemcpyAsync(dst[k], s->array(k), s->count * sizeof(double),
                                   hipMemcpyDeviceToHost, s->ctx->stream));
        }
        s->ctx->sync();
    });
}

int orr_system_refresh(orr_system* s, const orr_step_config* c) {
    return guarded([&] {
        require(s != nullptr, "null system");
        check_config(c);
        use_device(s->ctx);
        if (s->count > 0) field_eval(s, *c);
        s->ctx->sync();
    });
}

int orr_system_step(orr_system* s, const orr_step_config* c, double dt, uint64_t steps) {
    return guarded([&] {
        require(s != nullptr, "null system");
        check_config(c);
        use_device(s->ctx);
        if (s->count == 0) return;
        auto one_step = [&] {
            if (c->integrator == 2) {
                rk4_step(s, *c, dt);
            } else if (c->integrator == 1) {
                vv_substep(s, *c, kOuterWeight * dt);
                vv_substep(s, *c, kInnerWeight * dt);
                vv_substep(s, *c, kOuterWeight * dt);
            } else {
                vv_substep(s, *c, dt);
            }
        };
        // A whole step is captured once and replayed (one graph launch a
        // step instead of its kicks, drifts and evaluation launches), keyed by
        // everything the enqueued work depends on; with phase timing on it
        // runs eagerly, timed.
        std::vector<unsigned char> key;
        orr::key_add(key, s->state.ptr);
        orr::key_add(key, s->count);
        orr::key_add(key, c->integrator);
        orr::key_add(key, c->solver);
        orr::key_add(key, c->softening_squared);
        orr::key_add(key, c->gravitational_constant);
        orr::key_add(key, c->opening_angle);
        orr::key_add(key, c->leaf_capacity);
        orr::key_add(key, c->quadrupole);
        orr::key_add(key, dt);
        orr_ctx* ctx = s->ctx;
        orr::key_add(key, ctx->direct_targets_per_lane);
        orr::key_add(key, ctx->direct_refine);
        orr::key_add(key, ctx->direct_block_target);
        orr::key_add(key, ctx->direct_pair_mode);
        orr::key_add(key, ctx->direct_equal_masses);
        orr::key_add(key, ctx->direct_small_order);
        orr::key_add(key, ctx->walk_targets_per_lane);
        orr::key_add(key, ctx->walk_lanes);
        for (uint64_t k = 0; k < steps; ++k) orr::run_captured(ctx, s->step_graph, key, one_step, false);
        s->ctx->sync();
    });
}

int orr_system_diagnostics(orr_system* s, double e2, double G, double out[8]) {
    return guarded([&] {
        require(s != nullptr && out != nullptr, "null argument");
        use_device(s->ctx);
        orr::diagnostics(s->ctx, s->count, s->array(0), s->array(1), s->array(2), s->array(3),
                         s->array(4), s->array(5), s->array(9), e2, G, out);
    });
}

}  // extern "C"
