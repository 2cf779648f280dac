// Serving: a request loop with continuous batching (host code).
//
// Requests carry an initial state, an integrator, a timestep and a step count.
// The server admits queued requests in arrival order as the HBM budget allows
// (first fit: a request that does not fit lets later, smaller ones pass),
// keeps each admitted request's state on the device until it is done, and
// advances all of them in scheduling rounds:
//
//   1. admit what fits;
//   2. rebuild the batches whose membership changed: small requests of one
//      integrator (and one softening and G) share a batch, requests of
//      `alone_from` particles or more get their own; a continuing request's
//      state moves device to device, a new one is uploaded and its starting
//      diagnostics and the refresh are enqueued on the batch's stream;
//   3. enqueue a quantum of steps on every batch -- each batch on its own
//      streams, so the batches run concurrently -- sized so that the round
//      lasts about `round_ms` by the batch's last measured step time, and never
//      past the nearest finish of its members; wait for all;
//   4. retire the requests whose last quantum ran: their end diagnostics and
//      final state were enqueued behind that quantum on their batch's stream,
//      so they overlap the other batches' work; memory returns to the budget.
//
// A request's arithmetic does not depend on its batch (see batch.hip), nor on
// how its steps are split into quanta, so its result is the same alone and in
// any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "engine.hpp"

namespace {

// HBM charged per admitted request: the upper bound of what the request holds
// as one system of a batch (orr_batch_bytes_bound: state, tables, ordering,
// tree or direct scratch, diagnostics), a second copy of its state and
// diagnostics scratch for the moment a rebuild holds the old and the new batch
// at once, and a fixed 4 MiB for allocator rounding and small tables.
constexpr std::uint64_t kBytesPerRequest = 4u << 20;
constexpr double kMaxBudget = 288e9;

enum : int { kQueued = 0, kRunning = 1, kDone = 2 };

struct Request {
    orr_request spec{};
    std::vector<double> host[7];  // initial state while queued
    int state = kQueued;
    bool fresh = true;      // not yet on the device
    bool finishing = false;  // its last quantum, end diagnostics and copy are enqueued
    std::uint64_t steps_done = 0;
    double submitted = 0, admitted = 0, completed = 0;
    double* diag = nullptr;    // pinned: start diagnostics [0, 8), end [8, 16)
    double* result = nullptr;  // pinned: final x y z vx vy vz once done, until taken
    bool taken = false;
    std::uint64_t bytes = 0;
};

void release(Request& r) noexcept {
    if (r.diag != nullptr) (void)hipHostFree(r.diag);
    if (r.result != nullptr) (void)hipHostFree(r.result);
    r.diag = r.result = nullptr;
}

struct GroupKey {
    int integrator;
    double eps2, G;
    std::uint64_t alone;  // request id when the request runs alone, else 0
    bool operator<(const GroupKey& o) const {
        return std::tie(integrator, eps2, G, alone) < std::tie(o.integrator, o.eps2, o.G, o.alone);
    }
};

struct Group {
    std::vector<std::uint64_t> members;  // request ids; batch system k = members[k] after a build
    std::map<std::uint64_t, std::uint32_t> slot;  // id -> system index in `batch`
    orr_batch* batch = nullptr;
    bool dirty = false;
    double ms_per_step = 0;  // last measured, sharing the device with the other batches
    std::uint64_t particles = 0;
    std::uint64_t quantum = 0;
    hipEvent_t start = nullptr, stop = nullptr;  // around the quantum, on the batch's stream
};

void release_events(Group& g) noexcept {
    if (g.start != nullptr) (void)hipEventDestroy(g.start);
    if (g.stop != nullptr) (void)hipEventDestroy(g.stop);
    g.start = g.stop = nullptr;
}

void call(int status) {
    if (status != ORR_OK) orr::fail(status, orr_last_error());
}

}  // namespace

struct orr_server {
    orr_ctx* ctx = nullptr;
    orr_server_config cfg{};
    std::chrono::steady_clock::time_point t0;
    std::uint64_t next_id = 1;
    std::map<std::uint64_t, Request> requests;
    std::deque<std::uint64_t> queue;
    std::map<GroupKey, Group> groups;
    std::uint64_t resident = 0, peak = 0, budget = 0;
    std::uint64_t rounds = 0, builds = 0, particle_steps = 0, completed = 0;
    std::uint32_t fail_rebuilds = 0;  // testing aid: rebuilds still to fail
    double busy = 0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

GroupKey key_of(const orr_server* s, std::uint64_t id, const Request& r) {
    const bool alone = r.spec.count >= s->cfg.alone_from;
    return GroupKey{r.spec.integrator, r.spec.softening_squared, r.spec.gravitational_constant,
                    alone ? id : 0};
}

orr_batch_config batch_config(const orr_server* s, std::int32_t integrator) {
    orr_batch_config bc{};
    bc.integrator = integrator;
    bc.quadrupole = s->cfg.quadrupole;
    bc.opening_angle = s->cfg.opening_angle;
    bc.leaf_capacity = s->cfg.leaf_capacity;
    bc.tiles_per_item = s->cfg.tiles_per_item;
    bc.direct_below = s->cfg.direct_below;
    bc.use_graph = s->cfg.use_graph;
    return bc;
}

std::uint64_t request_bytes(const orr_server* s, std::uint64_t count) {
    const orr_batch_config bc = batch_config(s, 0);
    return orr::system_bytes_bound(count, -1, bc) + orr::grown_bytes(80 * count) +
           orr::grown_bytes(orr::diagnostics_scratch_bytes(count)) + kBytesPerRequest;
}

// Builds the group's batch from its members. Nothing the group or its
// requests hold changes until the new batch is complete: a failure on the way
// (an allocation, an upload, a device-to-device copy) leaves the old batch, the
// slots and every new request's host copy as they were, so a later round can
// try again.
void rebuild(orr_server* s, const GroupKey& key, Group& g) {
    orr_batch* old = g.batch;
    if (g.members.empty()) {
        g.batch = nullptr;
        g.slot.clear();
        g.particles = 0;
        g.dirty = false;
        if (old != nullptr) call(orr_batch_destroy(old));
        return;
    }
    std::vector<orr_batch_system> sys;
    std::uint64_t particles = 0;
    for (std::uint64_t id : g.members) {
        const orr_request& q = s->requests.at(id).spec;
        sys.push_back(orr_batch_system{q.count, -1, 0, q.softening_squared, q.gravitational_constant,
                                       q.timestep});
        particles += q.count;
    }
    const orr_batch_config bc = batch_config(s, key.integrator);
    orr_batch* nb = nullptr;
    call(orr_batch_create(s->ctx, sys.data(), static_cast<std::uint32_t>(sys.size()), &bc, &nb));
    std::map<std::uint64_t, std::uint32_t> slot;
    try {
        for (std::uint32_t k = 0; k < g.members.size(); ++k) {
            const std::uint64_t id = g.members[k];
            Request& r = s->requests.at(id);
            slot[id] = k;
            if (r.fresh) {
                call(orr_batch_upload(nb, k, r.host[0].data(), r.host[1].data(), r.host[2].data(),
                                      r.host[3].data(), r.host[4].data(), r.host[5].data(),
                                      r.host[6].data()));
                // Start diagnostics on the batch's stream, ahead of its steps.
                orr::batch_diagnostics_enqueue(nb, k, r.diag);
            } else {
                call(orr_batch_copy_system(nb, k, old, g.slot.at(id)));
            }
        }
        if (s->fail_rebuilds > 0) {
            --s->fail_rebuilds;
            orr::fail(ORR_ERR_RUNTIME, "injected failure of a batch rebuild (orr_server_config.fail_rebuilds)");
        }
    } catch (...) {
        (void)orr_batch_destroy(nb);
        throw;
    }
    // Commit: the new batch holds every member.
    g.batch = nb;
    g.slot = std::move(slot);
    for (std::uint64_t id : g.members) {
        Request& r = s->requests.at(id);
        if (!r.fresh) continue;
        for (auto& v : r.host) std::vector<double>().swap(v);
        r.fresh = false;
    }
    if (g.particles > 0 && g.ms_per_step > 0) {
        g.ms_per_step *= static_cast<double>(particles) / static_cast<double>(g.particles);
    }
    g.particles = particles;
    ++s->builds;
    // The copies are done: the old batch goes before the new one allocates its
    // scratch, so the two never hold scratch at once.
    if (old != nullptr) call(orr_batch_destroy(old));
    // Accelerations for the current positions (a continuing request's come
    // out as the bits it already had), enqueued before the steps.
    orr::batch_refresh_enqueue(nb);
    g.dirty = false;
}

void admit(orr_server* s) {
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        Request& r = s->requests.at(*it);
        if (s->resident + r.bytes > s->budget) {
            ++it;
            continue;
        }
        // Pinned host memory for the diagnostics and the final state, taken at
        // admission so that nothing allocates while batches run.
        if (r.diag == nullptr) ORR_HIP(hipHostMalloc(reinterpret_cast<void**>(&r.diag), 16 * sizeof(double)));
        if (r.result == nullptr) {
            ORR_HIP(hipHostMalloc(reinterpret_cast<void**>(&r.result), 6 * r.spec.count * sizeof(double)));
        }
        r.state = kRunning;
        r.admitted = s->now();
        s->resident += r.bytes;
        s->peak = std::max(s->peak, s->resident);
        Group& g = s->groups[key_of(s, *it, r)];
        g.members.push_back(*it);
        g.dirty = true;
        it = s->queue.erase(it);
    }
}

// After a round: requests whose last quantum ran (their end diagnostics and
// final state were enqueued behind it) are done.
void retire(orr_server* s, Group& g) {
    std::vector<std::uint64_t> keep;
    for (std::uint64_t id : g.members) {
        Request& r = s->requests.at(id);
        if (!r.finishing) {
            keep.push_back(id);
            continue;
        }
        r.finishing = false;
        r.state = kDone;
        r.completed = s->now();
        s->resident -= r.bytes;
        ++s->completed;
        g.dirty = true;
    }
    g.members = std::move(keep);
}

void round(orr_server* s) {
    const double start = s->now();
    // Groups that lost members last round still hold their memory in their
    // old batches: rebuild them before admitting into the budget they freed.
    if (!s->queue.empty()) {
        for (auto& [key, g] : s->groups) {
            if (g.dirty) rebuild(s, key, g);
        }
    }
    admit(s);
    for (auto& [key, g] : s->groups) {
        if (g.dirty) rebuild(s, key, g);
    }
    std::vector<Group*> running;
    const double target = s->cfg.round_ms;
    for (auto& [key, g] : s->groups) {
        if (g.members.empty()) continue;
        std::uint64_t remaining = ~std::uint64_t{0};
        for (std::uint64_t id : g.members) {
            const Request& r = s->requests.at(id);
            remaining = std::min(remaining, r.spec.steps - r.steps_done);
        }
        std::uint64_t q = 1;  // first round of a group: one step, to measure it
        if (g.ms_per_step > 0) {
            q = static_cast<std::uint64_t>(std::max(1.0, std::floor(target / g.ms_per_step)));
        }
        g.quantum = std::min(q, remaining);
        if (g.start == nullptr) ORR_HIP(hipEventCreate(&g.start));
        if (g.stop == nullptr) ORR_HIP(hipEventCreate(&g.stop));
        ORR_HIP(hipEventRecord(g.start, orr::batch_stream(g.batch)));
        orr::batch_step_enqueue(g.batch, g.quantum);
        ORR_HIP(hipEventRecord(g.stop, orr::batch_stream(g.batch)));
        // Members this quantum finishes: end diagnostics and final state behind it.
        for (std::uint64_t id : g.members) {
            Request& r = s->requests.at(id);
            if (r.spec.steps - r.steps_done != g.quantum) continue;
            const std::uint32_t k = g.slot.at(id);
            orr::batch_diagnostics_enqueue(g.batch, k, r.diag + 8);
            orr::batch_download_enqueue(g.batch, k, r.result);
            r.finishing = true;
        }
        running.push_back(&g);
    }
    for (Group* g : running) orr::batch_wait(g->batch);
    for (Group* g : running) {
        float ms = 0.f;
        ORR_HIP(hipEventElapsedTime(&ms, g->start, g->stop));
        g->ms_per_step = static_cast<double>(ms) / static_cast<double>(g->quantum);
        for (std::uint64_t id : g->members) {
            Request& r = s->requests.at(id);
            r.steps_done += g->quantum;
            s->particle_steps += r.spec.count * g->quantum;
        }
        retire(s, *g);
    }
    for (auto it = s->groups.begin(); it != s->groups.end();) {
        Group& g = it->second;
        if (g.members.empty()) {
            if (g.batch != nullptr) call(orr_batch_destroy(g.batch));
            release_events(g);
            it = s->groups.erase(it);
        } else {
            ++it;
        }
    }
    ++s->rounds;
    s->busy += s->now() - start;
}

void destroy(orr_server* s) noexcept {
    if (s == nullptr) return;
    for (auto& [key, g] : s->groups) {
        if (g.batch != nullptr) orr_batch_destroy(g.batch);
        release_events(g);
    }
    for (auto& [id, r] : s->requests) release(r);
    delete s;
}

Request& find(orr_server* s, std::uint64_t id) {
    auto it = s->requests.find(id);
    orr::require(it != s->requests.end(), "unknown request id");
    return it->second;
}

}  // namespace

using orr::guarded;
using orr::require;

extern "C" {

int orr_server_create(orr_ctx* ctx, const orr_server_config* config, orr_server** out) {
    return guarded([&] {
        require(ctx != nullptr && config != nullptr && out != nullptr, "null argument");
        *out = nullptr;
        ORR_HIP(hipSetDevice(ctx->device));
        std::unique_ptr<orr_server, void (*)(orr_server*)> s(new orr_server, destroy);
        s->ctx = ctx;
        s->cfg = *config;
        if (s->cfg.round_ms <= 0) s->cfg.round_ms = 50.0;
        if (s->cfg.direct_below == 0) s->cfg.direct_below = 4096;
        if (s->cfg.alone_from == 0) s->cfg.alone_from = 131072;
        if (!(s->cfg.opening_angle > 0)) s->cfg.opening_angle = 0.5;
        if (s->cfg.leaf_capacity == 0) s->cfg.leaf_capacity = 16;
        if (s->cfg.tiles_per_item == 0) s->cfg.tiles_per_item = 16;
        std::uint64_t budget = config->memory_budget;
        if (budget == 0) {
            std::size_t free_b = 0, total_b = 0;
            ORR_HIP(hipMemGetInfo(&free_b, &total_b));
            budget = static_cast<std::uint64_t>(std::min(kMaxBudget, 0.9 * static_cast<double>(free_b)));
        }
        s->budget = budget;
        s->fail_rebuilds = config->fail_rebuilds;
        s->t0 = std::chrono::steady_clock::now();
        *out = s.release();
    });
}

int orr_server_destroy(orr_server* s) {
    return guarded([&] { destroy(s); });
}

int orr_server_now(orr_server* s, double* seconds) {
    return guarded([&] {
        require(s != nullptr && seconds != nullptr, "null argument");
        *seconds = s->now();
    });
}

int orr_server_submit(orr_server* s, const orr_request* q, uint64_t* id) {
    return guarded([&] {
        require(s != nullptr && q != nullptr && id != nullptr, "null argument");
        require(q->count > 0, "a request needs at least one particle");
        require(q->count < (std::uint64_t{1} << 31), "a request holds fewer than 2^31 particles");
        require(q->integrator == 0 || q->integrator == 1, "integrator must be 0 (VV) or 1 (Yoshida-4)");
        require(q->steps > 0, "a request needs at least one step");
        require(q->softening_squared >= 0, "softening squared must be >= 0");
        require(std::isfinite(q->timestep), "timestep must be finite");
        const double* src[7] = {q->x, q->y, q->z, q->vx, q->vy, q->vz, q->m};
        for (const double* p : src) require(p != nullptr, "request arrays must not be null");
        Request r;
        r.spec = *q;
        r.spec.x = r.spec.y = r.spec.z = r.spec.vx = r.spec.vy = r.spec.vz = r.spec.m = nullptr;
        r.bytes = request_bytes(s, q->count);
        if (r.bytes > s->budget) orr::fail(ORR_ERR_OUT_OF_MEMORY, "request larger than the memory budget");
        for (int a = 0; a < 7; ++a) r.host[a].assign(src[a], src[a] + q->count);
        r.submitted = s->now();
        const std::uint64_t k = s->next_id++;
        s->requests.emplace(k, std::move(r));
        s->queue.push_back(k);
        *id = k;
    });
}

int orr_server_round(orr_server* s, uint32_t* pending) {
    return guarded([&] {
        require(s != nullptr, "null server");
        ORR_HIP(hipSetDevice(s->ctx->device));
        round(s);
        if (pending != nullptr) {
            std::uint64_t n = s->queue.size();
            for (const auto& [key, g] : s->groups) n += g.members.size();
            *pending = static_cast<uint32_t>(n);
        }
    });
}

int orr_server_status(orr_server* s, uint64_t id, orr_request_status* out) {
    return guarded([&] {
        require(s != nullptr && out != nullptr, "null argument");
        const Request& r = find(s, id);
        *out = orr_request_status{};
        out->state = r.state;
        out->steps_done = r.steps_done;
        out->submitted_s = r.submitted;
        out->admitted_s = r.admitted;
        out->completed_s = r.completed;
        out->charged_bytes = r.bytes;
        // Start diagnostics exist once a round has run since admission, end
        // diagnostics once the request is done.
        if (r.diag != nullptr && r.state != kQueued) std::memcpy(out->diagnostics_start, r.diag, 8 * sizeof(double));
        if (r.diag != nullptr && r.state == kDone) std::memcpy(out->diagnostics_end, r.diag + 8, 8 * sizeof(double));
    });
}

int orr_server_result(orr_server* s, uint64_t id, double* x, double* y, double* z, double* vx,
                      double* vy, double* vz) {
    return guarded([&] {
        require(s != nullptr, "null server");
        Request& r = find(s, id);
        require(r.state == kDone, "request not finished");
        require(!r.taken, "result already taken");
        double* dst[6] = {x, y, z, vx, vy, vz};
        for (int a = 0; a < 6; ++a) {
            if (dst[a] != nullptr) {
                std::memcpy(dst[a], r.result + static_cast<std::size_t>(a) * r.spec.count,
                            r.spec.count * sizeof(double));
            }
        }
        (void)hipHostFree(r.result);
        r.result = nullptr;
        r.taken = true;
    });
}

int orr_server_stats_get(orr_server* s, orr_server_stats* out) {
    return guarded([&] {
        require(s != nullptr && out != nullptr, "null argument");
        *out = orr_server_stats{};
        out->rounds = s->rounds;
        out->batch_builds = s->builds;
        out->particle_steps = s->particle_steps;
        out->completed = s->completed;
        out->resident_bytes = s->resident;
        out->peak_resident_bytes = s->peak;
        out->memory_budget = s->budget;
        std::uint64_t active = 0;
        for (const auto& [key, g] : s->groups) active += g.members.size();
        out->active = static_cast<uint32_t>(active);
        out->queued = static_cast<uint32_t>(s->queue.size());
        out->busy_s = s->busy;
    });
}

}  // extern "C"
