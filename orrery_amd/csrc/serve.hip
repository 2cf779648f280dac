// Serving: a request loop with continuous batching (host code).
//
// Requests carry an initial state, an integrator, a timestep and a step count.
// Admitted requests keep their state on the device until done. Small requests
// of one integrator (and one softening and G) share a batch; a request of
// `alone_from` particles or more is large and gets a batch of its own. Every
// group (one batch, on a pair of contexts -- streams and scratch -- that it
// keeps across rebuilds) advances on its own, concurrently with the others:
// there is no round barrier. One call of round():
//
//   1. retires the quanta that finished (polled by event, never waited for):
//      their members' step counts advance; requests whose last quantum ran
//      are done -- their end diagnostics and final state were enqueued behind
//      that quantum on their batch's stream -- and their memory returns to
//      the budget;
//   2. admits queued requests in arrival order as the HBM budget allows (first
//      fit: a small request that does not fit lets later, smaller ones pass).
//      Large requests run one at a time, in arrival order: a later large
//      request waits for the one ahead of it to finish. Small requests never
//      wait for a large one, and a large one is never stretched by another
//      large one sharing the device, so the longest requests finish no later
//      than first-in-first-out service would finish them, while the small ones
//      overtake;
//   3. on every group with no quantum in flight: rebuilds its batch if its
//      membership changed (a continuing request's state moves device to
//      device; a new one is uploaded, its starting diagnostics and the refresh
//      are enqueued on the batch's stream), then enqueues its next quantum of
//      steps, sized to about `round_ms` by the group's last measured step time
//      and never past the nearest finish of its members;
//   4. sleeps 0.1 ms when none of this found anything to do.
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
#include <thread>
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

// The two contexts a group's batches run on (direct and Barnes-Hut branch).
struct Lanes {
    orr_ctx* d = nullptr;
    orr_ctx* t = nullptr;
};

struct Group {
    std::vector<std::uint64_t> members;  // request ids; batch system k = members[k] after a build
    std::map<std::uint64_t, std::uint32_t> slot;  // id -> system index in `batch`
    orr_batch* batch = nullptr;
    Lanes lanes;
    bool dirty = false;
    bool in_flight = false;  // a quantum is enqueued and not yet retired
    double ms_per_step = 0;  // last measured, sharing the device with the other groups
    std::uint64_t particles = 0;
    std::uint64_t quantum = 0;
    // Around the quantum's steps (timing), and after everything the quantum
    // enqueued, finishing members' diagnostics and downloads included.
    hipEvent_t start = nullptr, stop = nullptr, done = nullptr;
};

void release_events(Group& g) noexcept {
    for (hipEvent_t* e : {&g.start, &g.stop, &g.done}) {
        if (*e != nullptr) (void)hipEventDestroy(*e);
        *e = nullptr;
    }
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
    std::uint64_t large_active = 0;  // the admitted large request, 0: none
    std::vector<Lanes> spare;        // contexts of groups that emptied, for new groups
    double busy = 0;                 // wall time with some quantum in flight
    double last_round = 0;
    bool was_busy = false;

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

Lanes take_lanes(orr_server* s) {
    if (!s->spare.empty()) {
        const Lanes l = s->spare.back();
        s->spare.pop_back();
        return l;
    }
    Lanes l;
    l.d = orr::make_context(s->ctx->device);
    try {
        l.t = orr::make_context(s->ctx->device);
    } catch (...) {
        orr::free_context(l.d);
        throw;
    }
    // The server context's kernel settings carry over.
    l.t->walk_targets_per_lane = s->ctx->walk_targets_per_lane;
    l.t->walk_lanes = s->ctx->walk_lanes;
    l.d->direct_refine = s->ctx->direct_refine;
    return l;
}

// A group that emptied: its streams go back to the pool, its scratch is freed
// (the budget charges scratch to the requests that used it).
void give_back_lanes(orr_server* s, Group& g) {
    if (g.lanes.d == nullptr) return;
    for (orr_ctx* c : {g.lanes.d, g.lanes.t}) {
        ORR_HIP(hipStreamSynchronize(c->stream));
        for (orr::Buffer* b : c->buffers()) b->release();
    }
    s->spare.push_back(g.lanes);
    g.lanes = Lanes{};
}

// Builds the group's batch from its members. Nothing the group or its
// requests hold changes until the new batch is complete: a failure on the way
// (an allocation, an upload, a device-to-device copy) leaves the old batch, the
// slots and every new request's host copy as they were, so a later round can
// try again. The old and the new batch share the group's streams.
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
    if (g.lanes.d == nullptr) g.lanes = take_lanes(s);
    std::vector<orr_batch_system> sys;
    std::uint64_t particles = 0;
    for (std::uint64_t id : g.members) {
        const orr_request& q = s->requests.at(id).spec;
        sys.push_back(orr_batch_system{q.count, -1, 0, q.softening_squared, q.gravitational_constant,
                                       q.timestep});
        particles += q.count;
    }
    const orr_batch_config bc = batch_config(s, key.integrator);
    orr_batch* nb = orr::batch_create_on(s->ctx, sys.data(), static_cast<std::uint32_t>(sys.size()),
                                         &bc, g.lanes.d, g.lanes.t);
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
    // The copies are done (orr_batch_copy_system waits for them).
    if (old != nullptr) call(orr_batch_destroy(old));
    // Accelerations for the current positions (a continuing request's come
    // out as the bits it already had), enqueued before the steps.
    orr::batch_refresh_enqueue(nb);
    g.dirty = false;
}

bool is_large(const orr_server* s, const Request& r) { return r.spec.count >= s->cfg.alone_from; }

void admit(orr_server* s) {
    // Large requests one at a time in arrival order: none passes an earlier
    // large request that is running or still waiting for memory.
    bool large_blocked = s->large_active != 0;
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        Request& r = s->requests.at(*it);
        const bool large = is_large(s, r);
        if (large && large_blocked) {
            ++it;
            continue;
        }
        if (s->resident + r.bytes > s->budget) {
            large_blocked = large_blocked || large;
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
        if (large) {
            s->large_active = *it;
            large_blocked = true;
        }
        Group& g = s->groups[key_of(s, *it, r)];
        g.members.push_back(*it);
        g.dirty = true;
        it = s->queue.erase(it);
    }
}

// A finished quantum: step counts advance, and the requests it finished (their
// end diagnostics and final state were enqueued behind it) are done.
void retire(orr_server* s, Group& g) {
    float ms = 0.f;
    ORR_HIP(hipEventElapsedTime(&ms, g.start, g.stop));
    g.ms_per_step = static_cast<double>(ms) / static_cast<double>(g.quantum);
    g.in_flight = false;
    std::vector<std::uint64_t> keep;
    for (std::uint64_t id : g.members) {
        Request& r = s->requests.at(id);
        // Admitted while the quantum ran: not in the batch yet, no steps.
        if (g.slot.count(id) == 0) {
            keep.push_back(id);
            continue;
        }
        r.steps_done += g.quantum;
        s->particle_steps += r.spec.count * g.quantum;
        if (!r.finishing) {
            keep.push_back(id);
            continue;
        }
        r.finishing = false;
        r.state = kDone;
        r.completed = s->now();
        s->resident -= r.bytes;
        ++s->completed;
        if (s->large_active == id) s->large_active = 0;
        g.dirty = true;
    }
    g.members = std::move(keep);
}

// The group's next quantum, enqueued on its batch's streams.
void launch_quantum(orr_server* s, Group& g) {
    std::uint64_t remaining = ~std::uint64_t{0};
    for (std::uint64_t id : g.members) {
        const Request& r = s->requests.at(id);
        remaining = std::min(remaining, r.spec.steps - r.steps_done);
    }
    std::uint64_t q = 1;  // first quantum of a group: one step, to measure it
    if (g.ms_per_step > 0) {
        q = static_cast<std::uint64_t>(std::max(1.0, std::floor(s->cfg.round_ms / g.ms_per_step)));
    }
    g.quantum = std::min(q, remaining);
    if (g.start == nullptr) ORR_HIP(hipEventCreate(&g.start));
    if (g.stop == nullptr) ORR_HIP(hipEventCreate(&g.stop));
    if (g.done == nullptr) ORR_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
    hipStream_t st = orr::batch_stream(g.batch);
    ORR_HIP(hipEventRecord(g.start, st));
    orr::batch_step_enqueue(g.batch, g.quantum);
    ORR_HIP(hipEventRecord(g.stop, st));
    // Members this quantum finishes: end diagnostics and final state behind it.
    for (std::uint64_t id : g.members) {
        Request& r = s->requests.at(id);
        if (r.spec.steps - r.steps_done != g.quantum) continue;
        const std::uint32_t k = g.slot.at(id);
        orr::batch_diagnostics_enqueue(g.batch, k, r.diag + 8);
        orr::batch_download_enqueue(g.batch, k, r.result);
        r.finishing = true;
    }
    ORR_HIP(hipEventRecord(g.done, st));
    g.in_flight = true;
}

void round(orr_server* s) {
    const double t = s->now();
    if (s->was_busy) s->busy += t - s->last_round;
    bool progressed = false;
    // 1. Quanta that finished.
    for (auto& [key, g] : s->groups) {
        if (!g.in_flight) continue;
        const hipError_t e = hipEventQuery(g.done);
        if (e == hipErrorNotReady) continue;
        ORR_HIP(e);
        retire(s, g);
        progressed = true;
    }
    // Groups that lost members still hold their memory in their old batches:
    // rebuild the idle ones before admitting into the budget they freed.
    if (!s->queue.empty()) {
        for (auto& [key, g] : s->groups) {
            if (g.dirty && !g.in_flight) rebuild(s, key, g);
        }
    }
    // 2. Admission.
    const std::size_t queued = s->queue.size();
    admit(s);
    progressed = progressed || s->queue.size() != queued;
    // 3. Idle groups: rebuild where membership changed, then the next quantum.
    for (auto& [key, g] : s->groups) {
        if (g.in_flight) continue;
        if (g.dirty) {
            rebuild(s, key, g);
            progressed = true;
        }
        if (g.members.empty()) continue;
        launch_quantum(s, g);
        progressed = true;
    }
    for (auto it = s->groups.begin(); it != s->groups.end();) {
        Group& g = it->second;
        if (g.members.empty() && !g.in_flight) {
            if (g.batch != nullptr) call(orr_batch_destroy(g.batch));
            g.batch = nullptr;
            give_back_lanes(s, g);
            release_events(g);
            it = s->groups.erase(it);
        } else {
            ++it;
        }
    }
    bool busy = false;
    for (const auto& [key, g] : s->groups) busy = busy || g.in_flight;
    // 4. Nothing to do until a quantum finishes: yield the host thread briefly.
    if (!progressed && busy) std::this_thread::sleep_for(std::chrono::microseconds(100));
    ++s->rounds;
    s->last_round = s->now();
    if (s->was_busy || busy) s->busy += s->last_round - t;
    s->was_busy = busy;
}

void destroy(orr_server* s) noexcept {
    if (s == nullptr) return;
    for (auto& [key, g] : s->groups) {
        if (g.batch != nullptr) orr_batch_destroy(g.batch);
        for (orr_ctx* c : {g.lanes.d, g.lanes.t}) orr::free_context(c);
        release_events(g);
    }
    for (const Lanes& l : s->spare) {
        orr::free_context(l.d);
        orr::free_context(l.t);
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
        // Start diagnostics exist once the request's first quantum has been
        // retired (they run ahead of it on its batch's stream), end
        // diagnostics once the request is done.
        if (r.diag != nullptr && (r.steps_done > 0 || r.state == kDone)) {
            std::memcpy(out->diagnostics_start, r.diag, 8 * sizeof(double));
        }
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
