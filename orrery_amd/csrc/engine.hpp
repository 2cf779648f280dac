// Internal interface of the MI355X force engine (not part of the C ABI).
//
// Errors inside the library are C++ exceptions carrying an ORR_* status; the
// extern "C" layer (capi.hip) catches them and returns the status, so no
// exception crosses the ABI (the reference's rule for its kernel translation
// units, include/orrery/solvers/cuda_kernels.hpp:96-111).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstddef>
#include <array>
#include <functional>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/orrery_hip.h"

namespace orr {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
    int status() const noexcept { return status_; }

private:
    int status_;
};

[[noreturn]] void fail(int status, const std::string& what);
void check_hip(hipError_t e, const char* what, const char* file, int line);

// The extern "C" layer's error channel: the message behind orr_last_error().
void set_last_error(const char* what);

// Runs body, mapping any exception to its status and message.
template<typename F> int guarded(F&& body) noexcept {
    try {
        body();
        return ORR_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.status();
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return ORR_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return ORR_ERR_RUNTIME;
    } catch (...) {
        set_last_error("unknown failure");
        return ORR_ERR_RUNTIME;
    }
}

inline void require(bool ok, const char* what) {
    if (!ok) fail(ORR_ERR_INVALID_ARGUMENT, what);
}

#define ORR_HIP(expr) ::orr::check_hip((expr), #expr, __FILE__, __LINE__)
#define ORR_LAUNCHED(what) ::orr::check_hip(hipGetLastError(), what, __FILE__, __LINE__)

// Count of device reallocations by any Buffer: a captured graph holds raw
// pointers, so it is valid only while this count is unchanged.
extern std::atomic<std::uint64_t> g_reallocations;

// A device allocation that only grows (the reference's solver pimpl rule,
// cuda_direct_solver.cpp:78-137).
// Bytes Buffer::reserve allocates for a request of `wanted` (a quarter more,
// so slowly growing systems do not reallocate every step).
std::size_t grown_bytes(std::size_t wanted);

struct Buffer {
    void* ptr = nullptr;
    std::size_t bytes = 0;
    template<typename T> T* as() const { return static_cast<T*>(ptr); }
    void reserve(std::size_t wanted);
    void release() noexcept;
};

enum Phase : int {
    kDirectPrepare = 0,
    kDirectKernel = 1,
    kDirectReduce = 2,
    kBbox = 3,
    kEncode = 4,
    kSort = 5,
    kGather = 6,
    kBuild = 7,
    kMoments = 8,
    kWalk = 9,
    kIntegrate = 10,
    kEnergy = 11,
};

}  // namespace orr

// The context owns its stream, its scratch and its timing events.
struct orr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // A second stream (default priority, made on first use) for work that can
    // overlap the main stream's within one evaluation, forked and joined by
    // the two events.
    hipStream_t aux = nullptr;
    hipEvent_t aux_fork = nullptr, aux_join = nullptr;
    int compute_units = 0;

    // Scratch, grown on demand.
    orr::Buffer direct_packed, direct_centres, direct_partials, direct_near;
    orr::Buffer order_aos;  // {x, y, z, m} per particle, index order (tree path)
    orr::Buffer order_codes, order_index, order_codes_alt, order_index_alt, order_hist, order_bbox;
    orr::Buffer tree_work;  // one arena: per-particle, per-node and counter arrays
    orr::Buffer diag_scratch;
    // Every scratch buffer above, for release and accounting.
    std::array<orr::Buffer*, 13> buffers() {
        return {&direct_packed,   &direct_centres,  &direct_partials, &direct_near, &order_aos,
                &order_codes,     &order_index,     &order_codes_alt, &order_index_alt,
                &order_hist,      &order_bbox,      &tree_work,       &diag_scratch};
    }

    // Direct-kernel configuration (orr_ctx_set_direct_config).
    int direct_targets_per_lane = 1;
    int direct_refine = 0;
    std::uint32_t direct_block_target = 32768;
    // 1: far tile pairs once for both tiles where that is faster and fits,
    // 2: one-sided, 3: pairs wherever they fit.
    int direct_pair_mode = 1;
    // 1: the pairs path's equal-mass form where the masses allow it, 2: never.
    int direct_equal_masses = 1;
    int direct_small_order = 1;  // 1: one-workgroup ordering up to kSmallOrder particles, 2: never
    int sort_mode = 1;  // Morton sort: 1 one-sweep passes, 2 histogram + scan + scatter per pass
    // Barnes-Hut walk configuration (orr_ctx_set_tree_config); 0 chooses by size.
    int walk_targets_per_lane = 0;
    int walk_lanes = 0;  // active lanes per walk wave: 32, 64, or 0 = by size

    // Host mirror of the last tree summary and the shape of its device layout.
    orr_tree_stats last_tree{};
    std::uint64_t tree_last_n = 0;
    std::size_t tree_last_bound = 0;
    double tree_last_theta = 0.5;
    bool tree_last_quadrupole = false;
    std::uint32_t tree_last_segments = 1;

    // Phase timing with hipEvents recorded on `stream`.
    bool timing = false;
    // Captured graphs allowed (false when the environment sets ORR_GRAPHS=0:
    // every sequence then runs eagerly, for tools that cannot follow graph
    // launches).
    bool graphs = true;
    struct Pending {
        int phase;
        hipEvent_t start, stop;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> free_events;
    orr_timings totals{};

    // Captured single-system evaluations (tree.hip, direct.hip; see
    // orr::run_captured), replayed while their arguments, kernel settings and
    // every buffer stay those they were captured with.
    struct Captured {
        std::vector<unsigned char> key;  // the arguments and settings, field by field
        std::uint64_t epoch = 0;         // g_reallocations when they last ran
        bool seen = false;               // this key has run once eagerly
        hipGraph_t graph = nullptr;      // kept for as long as its executable
        hipGraphExec_t exec = nullptr;
    };
    Captured tree_graph, direct_graph;
    // True while a run_captured sequence is being captured: a run_captured
    // reached from inside it then enqueues its work inline (into the same
    // capture) instead of launching or capturing a graph of its own.
    bool capturing = false;

    hipEvent_t take_event();
    void begin(int phase, hipEvent_t& start);
    void end(int phase, hipEvent_t start);
    // Synchronises the stream and folds finished event pairs into `totals`.
    void sync();
};

namespace orr {

// yoshida4.cpp:24-41 weights, evaluated in FP64 exactly as the reference's constexprs.
inline constexpr double kCubeRootOfTwo = 1.2599210498948731647672106;
inline constexpr double kOuterWeight = 1.0 / (2 - kCubeRootOfTwo);
inline constexpr double kInnerWeight = -kCubeRootOfTwo / (2 - kCubeRootOfTwo);

// context.hip: the context's second stream and its fork/join events, made on
// first use.
void ensure_aux(orr_ctx* ctx);

// context.hip: a context with its own stream (default priority); freeing
// synchronises it and releases every buffer.
orr_ctx* make_context(int device);
void free_context(orr_ctx* ctx) noexcept;

// RAII phase marker: records events around the launches it brackets.
class PhaseTimer {
public:
    // on == false records nothing whatever the context's timing setting (for
    // paths that must not create events or grow host tables).
    PhaseTimer(orr_ctx* ctx, int phase, bool on = true) : ctx_(on ? ctx : nullptr), phase_(phase) {
        if (ctx_ != nullptr) ctx_->begin(phase_, start_);
    }
    ~PhaseTimer() {
        if (ctx_ != nullptr) ctx_->end(phase_, start_);
    }
    PhaseTimer(const PhaseTimer&) = delete;
    PhaseTimer& operator=(const PhaseTimer&) = delete;

private:
    orr_ctx* ctx_;
    int phase_;
    hipEvent_t start_ = nullptr;
};

// ---- entry points implemented by the kernel translation units -------------

// A velocity kick v += s a carried out by a direct evaluation's fold right
// after it writes a: one launch fewer a step, the same bits as kick() (the
// multiply and the add rounded apart). vx == nullptr: no kick.
struct FusedKick {
    double* vx = nullptr;
    double* vy = nullptr;
    double* vz = nullptr;
    double s = 0.0;
};
__device__ __forceinline__ double kick_update(double v, double s, double a) {
#pragma clang fp contract(off)
    return v + s * a;
}

// direct.hip: enqueue the direct sum (no synchronisation).
void direct_enqueue(orr_ctx* ctx, const orr_direct_args& a, const std::uint32_t* order,
                    const FusedKick& kick = FusedKick{});

// direct.hip, batched: the direct systems of a batch as one tile space.
// Tables are device arrays; tile/item/slot meaning is given in direct.hip.
struct DirectBatchView {
    std::uint64_t count = 0;                 // particles (sorted positions 0..count-1)
    std::uint32_t systems = 0, tiles = 0, items = 0, slots = 0;
    const uint4* sys = nullptr;              // [systems] {first, count, first tile, splits}
    const uint4* tile = nullptr;             // [tiles] {first, valid, system, first slot}
    const uint4* item = nullptr;             // [items] {tile, source tile begin, end, slot}
    const std::uint32_t* seg_of = nullptr;   // [count] system of each sorted position
    const float* eps2 = nullptr;             // [systems] softening squared
    const double* G = nullptr;               // [systems] gravitational constant
    bool safe = false;                       // some system has zero softening
};
void direct_batch_enqueue(orr_ctx* ctx, const double* x, const double* y, const double* z,
                          const double* m, const std::uint32_t* order, const DirectBatchView& b,
                          double* ax, double* ay, double* az);

// Independent systems laid end to end in one set of arrays (a batch). All
// pointers are device arrays; count == 1 with null pointers is one system.
struct SegmentView {
    std::uint32_t count = 1;
    const std::uint32_t* start = nullptr;        // [count + 1] first particle of each system
    const std::uint32_t* seg_of = nullptr;       // [n] system of each particle
    const std::uint32_t* chunks = nullptr;       // [chunk_count][3] {system, begin, end}
    const std::uint32_t* chunk_first = nullptr;  // [count + 1] first chunk of each system
    std::uint32_t chunk_count = 0;
};

// morton.hip, segmented: one cube per system (cubes at ctx->order_bbox[4 * s]),
// order by (system, code, index). Systems keep their particle ranges. With
// masses given, the encode pass also leaves each particle's {x, y, z, m} in
// index order in ctx->order_aos (double4), so a later gather through the
// permutation reads one 32-byte record per particle instead of four scattered
// FP64 values.
void morton_enqueue_segments(orr_ctx* ctx, const double* x, const double* y, const double* z,
                             std::uint64_t n, const SegmentView& segments,
                             const double* m = nullptr, int first_digit = 0);

// Bytes the ordering stage reserves for n particles in `segments` systems
// (`chunks` bounding-box chunks when segmented), with or without the
// {x, y, z, m} records: the figures morton_enqueue_segments reserves with.
struct OrderSizes {
    std::uint32_t bbox_blocks, sort_blocks;
    std::size_t aos, bbox, codes, index, hist, total;
};
OrderSizes order_sizes(std::uint64_t n, std::uint32_t segments, std::uint32_t chunks, bool aos);

// The ordering stage's device buffers (sizes from order_sizes): codes and
// index twice (the radix passes alternate), the histogram table with its scan
// scratch, the cubes and bounding-box partials, the optional records.
struct OrderScratch {
    double* bbox = nullptr;
    double4* aos = nullptr;
    std::uint64_t* codes[2] = {nullptr, nullptr};
    std::uint32_t* index[2] = {nullptr, nullptr};
    std::uint32_t* hist = nullptr;
};
// morton_enqueue_segments on given buffers, phases timed when `timed`.
// Returns which of the two (codes, index) pairs holds the sorted order.
int morton_enqueue_into(orr_ctx* ctx, const double* x, const double* y, const double* z,
                        std::uint64_t n, const SegmentView& segments, const double* m,
                        int first_digit, const OrderScratch& buffers, bool timed);

// tree.hip: bytes tree_evaluate_segments reserves for n particles in `segments`
// systems at this leaf capacity (its node arrays are sized to the worst-case
// node count, n + 21 floor(n / (cap + 1)) + 22 per system).
std::size_t tree_bytes(std::uint64_t n, std::uint64_t leaf_capacity, bool quadrupole,
                       std::uint32_t segments);

// direct.hip: bytes direct_batch_enqueue reserves for `tiles` source tiles and
// `slots` partial slots.
std::size_t direct_batch_bytes(std::uint64_t tiles, std::uint64_t slots);

// batch.hip: an upper bound of the device bytes one system of n particles
// holds in a batch with this configuration (state, tables and every scratch
// buffer after any refresh, step, autotune or diagnostics). Bounds add: a
// batch of several systems holds at most the sum of theirs plus 4 KiB.
std::uint64_t system_bytes_bound(std::uint64_t n, int solver, const orr_batch_config& cfg);

// context.hip: a sequence of short launches with a fixed key (arguments and
// settings as bytes): the first run with a key is eager (and sizes every
// buffer), the second is captured as a hipGraph and launched, later ones replay
// it -- removing the host's launch gaps -- while the key and every buffer
// (g_reallocations) are unchanged. Events recorded inside a captured graph give
// no usable times, so with phase timing on it runs eagerly, unless `untimed`:
// then the sequence is replayed anyway and its phases go untimed.
void run_captured(orr_ctx* ctx, orr_ctx::Captured& g, std::vector<unsigned char> key,
                  const std::function<void()>& enqueue, bool untimed = false);
void drop_captured(orr_ctx::Captured& g) noexcept;

// Appends a value's bytes to a replay key.
template<typename T> void key_add(std::vector<unsigned char>& key, const T& v) {
    const auto* b = reinterpret_cast<const unsigned char*>(&v);
    key.insert(key.end(), b, b + sizeof(T));
}

// direct.hip: Morton order (when asked for; replayed as a graph where
// run_captured allows) and the direct sum.
void direct_evaluate(orr_ctx* ctx, const orr_direct_args& a, const FusedKick& kick = FusedKick{});
// direct.hip: the same sum with every buffer in the library's static scratch
// (no allocation, no graph, no phase timing; limits in direct.hip).
void direct_evaluate_static(orr_ctx* ctx, const orr_direct_args& a);
std::size_t direct_static_scratch_bytes();
std::uint64_t direct_static_max_count();

// context.hip: device bytes held by a context's scratch buffers.
std::size_t ctx_device_bytes(const orr_ctx* ctx);

// morton.hip: bounding cube, codes and stable radix sort from identity order.
// Leaves ascending codes and permutation in ctx->order_codes / order_index
// (uint64 / uint32). Returns nothing on the host; cube lands in ctx->order_bbox.
// first_digit > 0 skips the sort's low 8-bit digits: the order is then by the
// code's top bits only (stable within equal tops), enough for compact tiles.
void morton_enqueue(orr_ctx* ctx, const double* x, const double* y, const double* z,
                    std::uint64_t n, int first_digit = 0);
// morton_enqueue's index (and cube) for n <= kSmallOrder and first_digit >= 4 in
// one launch: the same permutation; ctx->order_codes is left as it was.
constexpr std::uint64_t kSmallOrder = 4096;
void morton_enqueue_small(orr_ctx* ctx, const double* x, const double* y, const double* z,
                          std::uint64_t n, int first_digit);
// Copies the cube {origin, size} computed by the last morton_enqueue (syncs).
void morton_cube(orr_ctx* ctx, double cube[4]);

// scan.hip: exclusive prefix sum of n uint32 (in == out allowed). scratch holds
// scan_scratch_words(n) words; *total_out (device, may be null) gets the sum.
std::size_t scan_scratch_words(std::uint64_t n);
void exclusive_scan_u32(orr_ctx* ctx, const std::uint32_t* in, std::uint32_t* out, std::uint64_t n,
                        std::uint32_t* scratch, std::uint32_t* total_out);

// tree.hip: order, build, moments and walk; fills stats (syncs).
void tree_evaluate(orr_ctx* ctx, const orr_tree_args& a, orr_tree_stats* stats);
// Segmented: one octree per system. seg_stats (device, may be null) receives
// {particle_particle, particle_cell, node_visits, 0} per system and must be
// zeroed by the caller; stats (may be null) gets the totals (syncs).
void tree_evaluate_segments(orr_ctx* ctx, const orr_tree_args& a, const SegmentView& segments,
                            unsigned long long* seg_stats, orr_tree_stats* stats);
void tree_copy_quadrupoles(orr_ctx* ctx, double* out);
// {particle_particle, particle_cell, node_visits, 0} per segment of the last
// segmented evaluation that kept its counts internally (seg_stats null); syncs.
void tree_copy_segment_counts(orr_ctx* ctx, std::uint64_t* out);
void tree_copy_nodes(orr_ctx* ctx, double* mass, double* cx, double* cy, double* cz,
                     double* acc, std::uint64_t* next, std::uint64_t* first,
                     std::uint64_t* count);

// capi.hip: a device-resident system's size and contiguous arrays (k = 0..9:
// x y z vx vy vz ax ay az m), copied `arrays` at a time from array `first`.
std::uint64_t system_count(const orr_system* s);
void system_download(orr_system* s, int first, int arrays, double* host);
void system_upload(orr_system* s, int first, int arrays, const double* host);

// batch.hip, for the serving loop: a batch on the caller's two contexts (its
// streams and scratch, which outlive the batch: a rebuilt batch reuses them).
orr_batch* batch_create_on(orr_ctx* owner, const orr_batch_system* systems, std::uint32_t count,
                           const orr_batch_config* config, orr_ctx* dctx, orr_ctx* tctx);
// Enqueue steps of a batch on its own streams without waiting (batches run
// concurrently), then wait for them.
void batch_step_enqueue(orr_batch* b, std::uint64_t steps);
void batch_wait(orr_batch* b);
// The stream every step of the batch starts and ends on.
hipStream_t batch_stream(orr_batch* b);
// Without waiting, on the batch's stream: accelerations for the current
// positions; a system's diagnostics into `out` (8 doubles); its x y z vx vy vz
// into `host` (6 x count doubles). Pinned host memory keeps the copies async.
void batch_refresh_enqueue(orr_batch* b);
void batch_diagnostics_enqueue(orr_batch* b, std::uint32_t system, double* out);
void batch_download_enqueue(orr_batch* b, std::uint32_t system, double* host);

// probes.hip: machine-ceiling probes; kind 0 fma, 1 rsq, 2 read, 3 copy.
void probe(orr_ctx* ctx, int kind, std::uint64_t size, double* rate, double* ms);

// integrate.hip: FP64 kick / drift / diagnostics on device-resident state.
void kick(orr_ctx* ctx, std::uint64_t n, double* vx, double* vy, double* vz, const double* ax,
          const double* ay, const double* az, double scale);
void drift(orr_ctx* ctx, std::uint64_t n, double* x, double* y, double* z, const double* vx,
           const double* vy, const double* vz, double scale);
// kick(v, a, kick_scale) then drift(x, v, drift_scale) in one pass, the same bits.
void kick_drift(orr_ctx* ctx, std::uint64_t n, double* x, double* y, double* z, double* vx, double* vy,
                double* vz, const double* ax, const double* ay, const double* az, double kick_scale,
                double drift_scale);
// The arrays of one RK4 step, per component: state, step start, stage sums.
struct Rk4Arrays {
    double* x[3];
    double* v[3];
    double* a[3];
    double* x0[3];
    double* v0[3];
    double* sx[3];
    double* sv[3];
};
// op 0 begin, 1 move to a stage (scale), 2 accumulate (weight), 3 combine (scale).
void rk4(orr_ctx* ctx, std::uint64_t n, const Rk4Arrays& arrays, int op, double scale);
void diagnostics(orr_ctx* ctx, std::uint64_t n, const double* x, const double* y,
                 const double* z, const double* vx, const double* vy, const double* vz,
                 const double* m, double softening_squared, double G, double out[8]);
// Enqueued on ctx->stream without waiting; `out` (8 doubles, best pinned host
// memory) holds {kinetic, potential, p, L} once the stream reaches it. The
// scratch (ctx->diag_scratch) grows only here: callers that enqueue several
// reserve diagnostics_scratch_bytes(largest n) first, so no growth frees memory
// that queued kernels still use.
std::size_t diagnostics_scratch_bytes(std::uint64_t n);
void diagnostics_enqueue(orr_ctx* ctx, std::uint64_t n, const double* x, const double* y,
                         const double* z, const double* vx, const double* vy, const double* vz,
                         const double* m, double softening_squared, double G, double* out);

}  // namespace orr
