/*
 * orrery_hip.h -- C ABI of the MI355X force engine for Orrery.
 *
 * The engine replaces the device halves of the reference's CUDA solvers
 * (src/solvers/cuda_direct_kernel.cu, cuda_tree_kernel.cu and the host halves
 * in cuda_direct_solver.cpp / cuda_tree_solver.cpp) behind the same
 * ForceSolver seam (include/orrery/solvers/force_solver.hpp:41-87). The C++
 * classes in orrery_amd/host/ and the Python classes in orrery_amd/solvers.py
 * are thin hosts over these entry points.
 *
 * Conventions (after the reference's launch boundary,
 * include/orrery/solvers/cuda_kernels.hpp:35-111, and its only C ABI,
 * wasm/orrery_wasm.h:48-203):
 *   - every entry point returns an int status, ORR_OK (0) on success; the
 *     message of the last failure on the calling thread is orr_last_error();
 *   - arguments are plain pointers and sizes bundled in POD structs; no C++
 *     or torch type crosses the boundary;
 *   - pointers named "device" are HIP device addresses of this context's
 *     device; every evaluation is synchronous: it returns once the results
 *     are in place (cuda_kernels.hpp:108-110);
 *   - particle arrays are structure-of-arrays FP64 (the reference's Real in
 *     its double build), accelerations are written, never accumulated
 *     (include/orrery/integrators/acceleration_field.hpp:48-49).
 */
#ifndef ORRERY_HIP_H
#define ORRERY_HIP_H

#include <stdint.h>

/* Exported from liborrery_hip.so; everything else in the library is hidden. */
#if defined(__GNUC__) || defined(__clang__)
#define ORR_API __attribute__((visibility("default")))
#else
#define ORR_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. Mirrors the exception kinds of the reference's CUDA host
 * (backend::CudaError for runtime failures, cuda_memory.hpp:129-153;
 * std::length_error for counts beyond 2^32, cuda_tree_solver.cpp:322-325). */
#define ORR_OK 0
#define ORR_ERR_INVALID_ARGUMENT 1
#define ORR_ERR_NO_DEVICE 2
#define ORR_ERR_RUNTIME 3
#define ORR_ERR_LENGTH 4
#define ORR_ERR_OUT_OF_MEMORY 5

typedef struct orr_ctx orr_ctx;
typedef struct orr_system orr_system;

/* ---- device discovery: replaces backend/cuda_device.hpp:61-209 ---------- */

typedef struct {
    char name[256];
    char arch[64];               /* e.g. "gfx950" */
    int32_t compute_units;
    int32_t wavefront_size;
    int32_t max_threads_per_block;
    int32_t clock_khz;
    uint64_t global_memory_bytes;
    uint64_t lds_per_block_bytes;
    uint64_t l2_bytes;
} orr_device_desc;

/* Number of HIP devices (0 when none; never an error for "no device"). */
ORR_API int orr_device_count(int32_t* count);
/* CudaDeviceDescription equivalent (cuda_device.hpp:61-100). */
ORR_API int orr_device_describe(int32_t device, orr_device_desc* out);
/* Message of the last failure on this thread (wasm/orrery_wasm.h:86-94). */
ORR_API const char* orr_last_error(void);

/* ---- context: one device, one stream, growing scratch buffers ----------- */

/* Replaces CudaDirectSolver::try_create / CudaTreeSolver::try_create device
 * set-up (cuda_direct_solver.cpp:140-164): ORR_ERR_NO_DEVICE when there is
 * none, so hosts can return null and fall back (sim/assembly.cpp:253-271). */
ORR_API int orr_ctx_create(int32_t device, orr_ctx** out);
ORR_API int orr_ctx_destroy(orr_ctx* ctx);

/* Device-memory plumbing for hosts that own no allocator
 * (backend/cuda_memory.hpp:77-348: CudaArray and its copies). */
ORR_API int orr_malloc(orr_ctx* ctx, uint64_t bytes, void** device_out);
ORR_API int orr_free(orr_ctx* ctx, void* device);
ORR_API int orr_upload(orr_ctx* ctx, void* device_dst, const void* host_src, uint64_t bytes);
ORR_API int orr_download(orr_ctx* ctx, void* host_dst, const void* device_src, uint64_t bytes);
ORR_API int orr_memset(orr_ctx* ctx, void* device_dst, int32_t value, uint64_t bytes);

/* ---- direct summation --------------------------------------------------- */

/* Replaces CudaDirectArguments + launch_cuda_direct (cuda_kernels.hpp:65-111;
 * cuda_direct_kernel.cu:87-214). a_i = G sum_{j != i} m_j d_ij / (r^2 + eps^2)^(3/2),
 * the law of core/softening.hpp:80-100. FP32 pairwise terms on tile-relative
 * coordinates, per-tile partial sums Kahan-added, FP64 result. */
typedef struct {
    const double* x;      /* device, count */
    const double* y;
    const double* z;
    const double* m;
    double* ax;           /* device, count; written */
    double* ay;
    double* az;
    uint64_t count;
    double softening_squared;
    double gravitational_constant;
    int32_t spatial_order;  /* 1: Morton-order the tiles first (default), 0: index order */
} orr_direct_args;
ORR_API int orr_direct_eval(orr_ctx* ctx, const orr_direct_args* args);

/* No counterpart in the reference: orr_direct_eval for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * hosts that forbid hipMalloc). Every buffer lives in one static device array
 * of 16 MiB per GPU, held from the first call on that GPU until the process
 * ends; nothing is allocated, no graph is built, phases are not timed. Same
 * arguments and result as orr_direct_eval, with these limits: count at most
 * 32,768 (ORR_ERR_LENGTH beyond); always the one-sided kernel; where the
 * one-sided kernel's split partials would not fit, fewer splits (the same
 * terms in fewer Kahan-summed groups, so not the same bits as orr_direct_eval;
 * the same bits whenever they fit and orr_direct_eval takes the one-sided
 * kernel). Calls that overlap on one GPU share the array and race. */
ORR_API int orr_direct_eval_static(orr_ctx* ctx, const orr_direct_args* args);
/* The limits above: the largest count and the scratch bytes per GPU. */
ORR_API int orr_direct_static_limits(uint64_t* max_count, uint64_t* scratch_bytes);

/* Launch configuration of the direct kernel (no reference counterpart: the
 * CUDA solver fixes its block size, cuda_direct_solver.cpp:46-63).
 * targets_per_lane in {1, 2, 4}; refine = 1 applies a Newton step to the
 * hardware reciprocal square root of the cubed argument (off by default: the
 * pair terms' FP32 rounding, the close pairs' double-float terms and the
 * compensated sums keep the force within 1e-6 relative without it); block_target is the number of thread blocks the
 * source range is split to reach (default 32768). pair_mode 1 (default)
 * evaluates each far pair of 64-particle tiles once and applies it to both
 * tiles where that is faster and fits: from 704 tiles (44,993 particles) up to
 * 856,064 particles (3 doubles per particle and later block of 8 tiles within
 * 16 GiB); pair_mode 3 does so wherever it fits (from 4 blocks of 8 tiles,
 * 1,537 particles); pair_mode 2 always runs the one-sided kernel, to which
 * targets_per_lane and block_target apply. equal_masses 1 (default) lets the
 * pairs path take its equal-mass form when every particle has one mass (the
 * mass leaves the pair term and multiplies the folded sums: two multiplies
 * fewer per pair; the same bits when that mass is a power of two), 2 never.
 * small_order 1 (default) orders a system of at most 4096 particles in one
 * workgroup (one launch instead of some twenty; the same permutation, so the
 * same bits), 2 never. targets_per_lane 0, refine -1, block_target 0,
 * pair_mode 0, equal_masses 0 and small_order 0 keep the current value. */
typedef struct {
    int32_t targets_per_lane;
    int32_t refine;
    uint32_t block_target;
    int32_t pair_mode;
    int32_t equal_masses;
    int32_t small_order;
} orr_direct_config;
ORR_API int orr_ctx_set_direct_config(orr_ctx* ctx, const orr_direct_config* config);
ORR_API int orr_ctx_get_direct_config(orr_ctx* ctx, orr_direct_config* config);

/* ---- Barnes-Hut --------------------------------------------------------- */

/* Replaces CudaTreeArguments + launch_cuda_tree (cuda_kernels.hpp:137-229) and
 * moves the reference's host ordering, sort and build (cuda_tree_solver.cpp:306-446)
 * onto the device. TreeParameters of solvers/octree.hpp:147-199; theta is
 * clamped to [0, 1] and leaf_capacity to >= 1 (octree.cpp:189-194). */
typedef struct {
    const double* x;      /* device, count */
    const double* y;
    const double* z;
    const double* m;
    double* ax;           /* device, count; written */
    double* ay;
    double* az;
    uint64_t count;
    double softening_squared;
    double gravitational_constant;
    double opening_angle;
    uint64_t leaf_capacity;
    int32_t quadrupole;
} orr_tree_args;

/* Interaction counts (solvers/interaction_count.hpp:37-76) and a host-readable
 * tree summary (Octree::leaf_count / depth, octree.hpp:229-236). */
typedef struct {
    uint64_t particle_particle;
    uint64_t particle_cell;
    uint64_t node_visits;
    uint64_t node_count;
    uint64_t leaf_count;
    uint32_t depth;
    uint32_t reserved;
    uint64_t wave_steps;     /* walk lane-target slots / 64, summed over waves */
} orr_tree_stats;
ORR_API int orr_tree_eval(orr_ctx* ctx, const orr_tree_args* args, orr_tree_stats* stats);

/* Walk configuration (no reference counterpart; cuda_tree_solver.cpp fixes its
 * traversal): targets_per_lane 1 or 2 -- how many targets each lane of a
 * wave walks together -- or -1 to choose by size (2 for one system of 2^20 or
 * more particles, else 1; the default); lanes 64 or 32 -- how many lanes of a
 * wave hold targets -- or -1 to choose by size (32 for a launch of at most
 * 2048 full waves, two per SIMD, else 64; the default). Results are
 * bit-identical either way; only the speed differs. sort_mode chooses the
 * Morton sort of every path (the tree's, the batch's, the direct ordering of
 * more than 4096 particles): 1 (the default) one launch per radix pass after
 * one launch that counts every pass's digits, 2 histogram, scan and scatter
 * launches per pass; the same permutation either way. Zero keeps the current
 * value; the getter reports -1 while a choice is by size. */
typedef struct {
    int32_t targets_per_lane;
    int32_t lanes;
    int32_t sort_mode;
    int32_t reserved;
} orr_tree_config;
ORR_API int orr_ctx_set_tree_config(orr_ctx* ctx, const orr_tree_config* config);
ORR_API int orr_ctx_get_tree_config(orr_ctx* ctx, orr_tree_config* config);

/* Morton ordering on the device: MortonOrdering::build (solvers/morton.cpp:180-197).
 * Writes the bounding cube {origin x, y, z, size} to host cube_out and the
 * ascending (code, index) keys to device arrays of `count` entries. */
ORR_API int orr_morton_order(orr_ctx* ctx, const double* x, const double* y, const double* z,
                     uint64_t count, uint64_t* codes_out, uint32_t* index_out,
                     double cube_out[4]);

/* Tree nodes of the last orr_tree_eval in DFS order (TreeNode, octree.hpp:106-139),
 * copied to host arrays of node_count entries; any pointer may be NULL. */
ORR_API int orr_tree_nodes(orr_ctx* ctx, double* mass, double* com_x, double* com_y, double* com_z,
                   double* acceptance_radius_squared, uint64_t* next,
                   uint64_t* first_particle, uint64_t* particle_count);

/* Quadrupole moments of the last orr_tree_eval run with quadrupole = 1, six
 * per node {xx, xy, xz, yy, yz, zz} (Quadrupole, octree.hpp:89-96), copied to
 * a host array of 6 * node_count doubles. */
ORR_API int orr_tree_quadrupoles(orr_ctx* ctx, double* quadrupoles);

/* ---- machine-ceiling probes ----------------------------------------------- */

/* Replaces the reference's CUDA probes (benchmarks/cuda_probes.cu:15-117:
 * fma_probe, div_sqrt_probe, read_probe). kind: 0 FP32 FMA chains (size =
 * iterations; rate in flop/s), 1 v_rsq_f32 chains (iterations; rsq/s),
 * 2 HBM read (size = bytes; bytes/s), 3 HBM copy (bytes copied; bytes read +
 * written per second), 4 one softened-gravity interaction's instruction mix
 * without memory (eleven FMAs, one add, one v_rsq_f32; size = iterations of
 * four; rate in interactions/s): the issue-rate ceiling of the direct kernels. */
#define ORR_PROBE_FMA 0
#define ORR_PROBE_RSQ 1
#define ORR_PROBE_READ 2
#define ORR_PROBE_COPY 3
#define ORR_PROBE_MIX 4
ORR_API int orr_probe(orr_ctx* ctx, int32_t kind, uint64_t size, double* rate_out,
                      double* milliseconds_out);

/* ---- phase timings (hipEvents on the context stream) -------------------- */

/* Phases: 0 direct prepare, 1 direct kernel, 2 direct reduce, 3 bbox, 4 encode,
 * 5 sort, 6 gather, 7 build, 8 moments, 9 walk, 10 integrate, 11 energy.
 * Enabled per context; the reference's EvaluationTimings / CudaEvaluationTimings
 * (barnes_hut_solver.hpp:92-105; cuda_direct_solver.hpp:85-...) carry the same
 * information. Milliseconds summed since the last reset, with launch counts. */
#define ORR_PHASES 16
typedef struct {
    double milliseconds[ORR_PHASES];
    uint64_t launches[ORR_PHASES];
    /* Captured evaluations replayed as one hipGraph launch since the last
     * reset (counted whether or not timing is on; with timing on evaluations
     * run eagerly so that their phases can be timed). */
    uint64_t graph_launches;
    uint64_t reserved[3];
} orr_timings;
ORR_API int orr_ctx_set_timing(orr_ctx* ctx, int32_t enabled);
ORR_API int orr_ctx_timings(orr_ctx* ctx, orr_timings* out);
ORR_API int orr_ctx_reset_timings(orr_ctx* ctx);

/* ---- device-resident systems: state stays in HBM across steps ----------- */

/* Integrators: 0 velocity Verlet (integrators/velocity_verlet.cpp:20-72),
 * 1 Yoshida-4 (integrators/yoshida4.cpp:24-59), 2 classical Runge-Kutta 4
 * (integrators/runge_kutta4.cpp:64-139; four evaluations per step, twelve more
 * FP64 arrays of stage state). Solvers: 0 direct, 1 Barnes-Hut.
 * FP64 kick and drift on the device; the field is the FP32-term engine above. */
typedef struct {
    int32_t integrator;
    int32_t solver;
    double softening_squared;
    double gravitational_constant;
    double opening_angle;
    uint64_t leaf_capacity;
    int32_t quadrupole;
    int32_t reserved;
} orr_step_config;

ORR_API int orr_system_create(orr_ctx* ctx, uint64_t count, orr_system** out);
ORR_API int orr_system_destroy(orr_system* sys);
/* Number of particles (ParticleData::size, core/particle_data.hpp). */
ORR_API int orr_system_count(orr_system* sys, uint64_t* count);
/* Host arrays of `count` FP64 values each. */
ORR_API int orr_system_upload(orr_system* sys, const double* x, const double* y, const double* z,
                      const double* vx, const double* vy, const double* vz, const double* m);
ORR_API int orr_system_download(orr_system* sys, double* x, double* y, double* z, double* vx,
                        double* vy, double* vz, double* ax, double* ay, double* az);
/* refresh_accelerations (integrators/integrator.hpp:51-53). */
ORR_API int orr_system_refresh(orr_system* sys, const orr_step_config* config);
ORR_API int orr_system_step(orr_system* sys, const orr_step_config* config, double timestep,
                    uint64_t steps);
/* measure_diagnostics (core/diagnostics.cpp:179-200) in FP64 on the device:
 * out = {kinetic, potential, px, py, pz, lx, ly, lz}. */
ORR_API int orr_system_diagnostics(orr_system* sys, double softening_squared,
                           double gravitational_constant, double out[8]);

/* ---- checkpoint and trajectory files from device-resident state ----------
 *
 * The reference's byte formats (docs/formats/checkpoint.md and
 * docs/formats/trajectory.md), so `orrery resume`, `orrery inspect` and the web
 * client read what the engine writes. */

/* sim/checkpoint.cpp:86-131: magic ORRERYCK, version 1, flags 0 (FP64), step,
 * N, configuration text (u32 length + bytes, stored as given), then masses and
 * x y z, vx vy vz, ax ay az, then the FNV-1a-64 checksum of all of it; written
 * to "<path>.partial" and renamed over `path`. */
ORR_API int orr_checkpoint_write(orr_system* sys, const char* path, uint64_t step,
                                 const char* configuration, uint64_t configuration_length);
/* sim/checkpoint.cpp:133-194: refuses a wrong magic, version, reserved or
 * single-precision flag, short file or checksum mismatch (ORR_ERR_INVALID_ARGUMENT);
 * else creates a system on ctx holding the stored state, accelerations
 * included, so a resumed run continues bit for bit. The configuration text is
 * copied NUL-terminated into `configuration` (capacity bytes, may be null);
 * parsing it is the caller's. */
ORR_API int orr_checkpoint_read(orr_ctx* ctx, const char* path, orr_system** out, uint64_t* step,
                                char* configuration, uint64_t capacity,
                                uint64_t* configuration_length);

/* sim/trajectory.cpp:38-101: the header (magic ORRERYTJ, version 1, flags with
 * bit 1 = frames carry velocities, N, timestep, masses, checksum) is written at
 * open; each append writes one frame (step, step x timestep, x y z [vx vy vz],
 * checksum) from the system's current device state. */
typedef struct orr_trajectory orr_trajectory;
ORR_API int orr_trajectory_open(orr_system* sys, const char* path, double timestep,
                                int32_t with_velocities, orr_trajectory** out);
ORR_API int orr_trajectory_append(orr_trajectory* trajectory, orr_system* sys, uint64_t step);
ORR_API int orr_trajectory_close(orr_trajectory* trajectory);

/* ---- batches: many independent systems of uneven size, stepped together --
 *
 * No counterpart in the reference, which steps one Simulation at a time
 * (sim/simulation.cpp:26-64 over integrators/velocity_verlet.cpp:59-72). A batch
 * is the device-resident system above for many systems at once: their FP64
 * state lies end to end in HBM, and one step of the whole batch is one
 * sequence of launches whatever the number of systems -- the direct systems
 * as one persistent work-queue kernel, the Barnes-Hut systems as one
 * segmented Morton sort, octree build and walk -- replayed as a hipGraph.
 * A system's result does not depend on the other systems in the batch.
 * One integrator per batch; softening, G and the timestep per system. */
typedef struct orr_batch orr_batch;

typedef struct {
    uint64_t count;                  /* particles (0 allowed: the system is inert) */
    int32_t solver;                  /* 0 direct, 1 Barnes-Hut, -1 by size (direct_below) */
    int32_t reserved;
    double softening_squared;        /* Barnes-Hut systems of one batch share it and G */
    double gravitational_constant;
    double timestep;
} orr_batch_system;

typedef struct {
    int32_t integrator;              /* 0 velocity Verlet, 1 Yoshida-4 */
    int32_t quadrupole;              /* Barnes-Hut systems */
    double opening_angle;
    uint32_t leaf_capacity;          /* 0: chosen by orr_batch_autotune (else 16) */
    uint32_t tiles_per_item;         /* direct work-item length in 64-source tiles; 0: tuned */
    uint32_t direct_below;           /* solver -1: direct when count < direct_below */
    int32_t use_graph;               /* capture one step as a hipGraph and replay it */
} orr_batch_config;

typedef struct {
    uint32_t systems, direct_systems, tree_systems;
    uint32_t tiles_per_item, leaf_capacity;
    uint32_t direct_tiles, direct_items;
    int32_t graph;                   /* a captured step is in use */
    uint64_t particles, direct_particles, tree_particles;
    uint64_t direct_pairs;           /* sum of N(N-1) over direct systems, per evaluation */
    uint64_t device_bytes;           /* device memory the batch holds now */
} orr_batch_info;

ORR_API int orr_batch_create(orr_ctx* ctx, const orr_batch_system* systems, uint32_t count,
                             const orr_batch_config* config, orr_batch** out);
ORR_API int orr_batch_destroy(orr_batch* batch);
ORR_API int orr_batch_info_get(orr_batch* batch, orr_batch_info* out);
/* No reference counterpart (the reference runs one system in host memory): an
 * upper bound of the device bytes a batch of these systems holds at its
 * largest -- state, tables and every scratch buffer after any refresh, step,
 * autotune or diagnostics -- for callers that plan device memory, such as the
 * serving loop's admission. Bounds of disjoint sets of systems add. */
ORR_API int orr_batch_bytes_bound(const orr_batch_system* systems, uint32_t count,
                                  const orr_batch_config* config, uint64_t* bytes_out);
/* Host arrays of the system's `count` FP64 values each (null: left as is). */
ORR_API int orr_batch_upload(orr_batch* batch, uint32_t system, const double* x, const double* y,
                             const double* z, const double* vx, const double* vy,
                             const double* vz, const double* m);
ORR_API int orr_batch_download(orr_batch* batch, uint32_t system, double* x, double* y, double* z,
                               double* vx, double* vy, double* vz, double* ax, double* ay,
                               double* az);
/* Device-to-device copy of one system's whole state (positions, velocities,
 * accelerations, masses) between batches; counts must match. */
ORR_API int orr_batch_copy_system(orr_batch* dst, uint32_t dst_system, orr_batch* src,
                                  uint32_t src_system);
/* Accelerations for the current positions of every system. */
ORR_API int orr_batch_refresh(orr_batch* batch);
/* `steps` steps of every system (synchronous). */
ORR_API int orr_batch_step(orr_batch* batch, uint64_t steps);
/* Times candidate tiles_per_item and leaf_capacity values on the batch's
 * current state and keeps the fastest (for fields left 0 in the config). */
ORR_API int orr_batch_autotune(orr_batch* batch);
ORR_API int orr_batch_diagnostics(orr_batch* batch, uint32_t system, double out[8]);
/* Interactions of the last evaluation of one system: particle_particle and
 * particle_cell (direct: N(N-1) and 0). */
ORR_API int orr_batch_counts(orr_batch* batch, uint32_t system, uint64_t out[2]);

/* ---- serving: a request loop with continuous batching --------------------
 *
 * No counterpart in the reference, which serves one Simulation per process
 * (its web service runs the wasm build per client, wasm/orrery_wasm.h:113-200).
 * Requests (an initial state, an integrator and a step count) are admitted
 * against an HBM budget as earlier ones finish; admitted requests stay on the
 * device until done and are stepped in batches (small systems of one
 * integrator together, large ones -- alone_from particles or more -- alone,
 * one large request at a time in arrival order), every batch advancing on its
 * own streams in quanta of steps, concurrently with the others. A request's
 * result does not depend on what else runs. */
typedef struct orr_server orr_server;

typedef struct {
    uint64_t memory_budget;          /* bytes of HBM for admitted requests; 0: the device's
                                        free memory at creation, at most 288 GB */
    double round_ms;                 /* target duration of one batch's quantum of steps (0: 50 ms) */
    uint32_t direct_below;           /* direct solver below this many particles (0: 4096) */
    uint32_t alone_from;             /* requests this large get a batch of their own (0: 131072) */
    double opening_angle;            /* Barnes-Hut requests (0: 0.5) */
    uint32_t leaf_capacity;          /* 0: 16 */
    uint32_t tiles_per_item;         /* 0: 16 */
    int32_t quadrupole;
    int32_t use_graph;
    /* Testing aid for the recovery path: the first `fail_rebuilds` batch
     * rebuilds fail after their uploads and copies, as a failed allocation or
     * copy would (orr_server_round returns ORR_ERR_RUNTIME); the requests and
     * the old batch are kept, and a later round tries again. 0 in service. */
    uint32_t fail_rebuilds;
    uint32_t reserved;
} orr_server_config;

typedef struct {
    uint64_t count;
    int32_t integrator;              /* 0 velocity Verlet, 1 Yoshida-4 */
    int32_t reserved;
    double softening_squared;
    double gravitational_constant;
    double timestep;
    uint64_t steps;
    /* Host arrays of `count` values, copied by orr_server_submit. */
    const double *x, *y, *z, *vx, *vy, *vz, *m;
} orr_request;

typedef struct {
    int32_t state;                   /* 0 queued, 1 running, 2 done */
    int32_t reserved;
    uint64_t steps_done;
    double submitted_s, admitted_s, completed_s;   /* server clock, seconds */
    double diagnostics_start[8];     /* at admission: {kinetic, potential, p, L} */
    double diagnostics_end[8];       /* at completion */
    uint64_t charged_bytes;          /* device memory the request holds against the budget */
} orr_request_status;

typedef struct {
    uint64_t rounds, batch_builds, particle_steps, completed;
    uint64_t resident_bytes, peak_resident_bytes, memory_budget;
    uint32_t active, queued;
    double busy_s;                   /* wall time with some batch's quantum in flight */
} orr_server_stats;

ORR_API int orr_server_create(orr_ctx* ctx, const orr_server_config* config, orr_server** out);
ORR_API int orr_server_destroy(orr_server* server);
/* Seconds since the server was created (the clock of the status times). */
ORR_API int orr_server_now(orr_server* server, double* seconds);
ORR_API int orr_server_submit(orr_server* server, const orr_request* request, uint64_t* id);
/* One pass of the loop: retire the quanta that finished (and the requests
 * they finished), admit, (re)batch, and start the next quantum on every batch
 * that has none in flight; sleeps 0.1 ms when there was nothing to do. It
 * never waits for a quantum in flight; rebuilding a batch (its copies) and
 * capturing a new batch's step wait for that batch's own streams only. Call it
 * repeatedly (submitting between calls). `pending` gets the number of
 * requests queued or running after it. */
ORR_API int orr_server_round(orr_server* server, uint32_t* pending);
ORR_API int orr_server_status(orr_server* server, uint64_t id, orr_request_status* out);
/* Final state of a finished request (host arrays, null: skipped); then its
 * host copy is released. */
ORR_API int orr_server_result(orr_server* server, uint64_t id, double* x, double* y, double* z,
                              double* vx, double* vy, double* vz);
ORR_API int orr_server_stats_get(orr_server* server, orr_server_stats* out);

/* ---- seeded initial conditions (host only; no device is touched) -------- */

/* The reference's Plummer sphere for a seed, bit for bit: replaces
 * make_plummer_sphere (src/initial_conditions/plummer.cpp:91-128) over
 * RandomSource (src/core/random.cpp:12-47; std::mt19937_64, top 53 bits) and
 * move_to_centre_of_mass_frame, with the position direction drawn before the
 * velocity direction (the Clang order of plummer.cpp:122). The input side of
 * the wasm ABI's orrery_create (wasm/orrery_wasm.h:113), which builds its
 * initial conditions inside the library. Writes `count` values into each host
 * array. ORR_ERR_INVALID_ARGUMENT unless total_mass > 0, scale_radius > 0 and
 * 0 < mass_fraction_cutoff < 1 (plummer.cpp:20-36); count 0 writes nothing. */
ORR_API int orr_plummer_sample(uint64_t count, double total_mass, double scale_radius,
                               double mass_fraction_cutoff, uint64_t seed, double* x, double* y,
                               double* z, double* vx, double* vy, double* vz, double* m);

#ifdef __cplusplus
}
#endif

#endif
