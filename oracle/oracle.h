/*
 * CPU oracle for the MI355X force engine -- TEST INFRASTRUCTURE ONLY.
 *
 * A restatement, in C++20 with a C ABI for ctypes, of the reference CPU
 * algorithms that the device path must agree with. Nothing in the product
 * (orrery_amd/, include/) links or calls this library: only tests/, the
 * smoke check and bench.py's CPU-baseline leg load it, as the checker.
 *
 * Each function names the reference file:line it follows. Parity status of
 * every function is listed in the header comment of oracle.cpp.
 *
 * Conventions: arrays are structure-of-arrays FP64 (x, y, z, m), counts are
 * uint64_t, returns are 0 on success and non-zero on a usage error.
 */
#ifndef ORRERY_ORACLE_H
#define ORRERY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random source and initial conditions ------------------------------ */

/* core/random.cpp:12-47: mt19937_64, 53-bit uniform, (height, azimuth) unit vector.
 * Draws `count` uniforms into `out` (for known-answer tests of the engine). */
int orc_random_uniform(uint64_t seed, uint64_t count, double* out);

/* initial_conditions/plummer.cpp:43-128 with the left-to-right draw order of
 * plummer.cpp:122 (position direction first, then velocity direction). */
int orc_plummer(uint64_t count, double total_mass, double scale_radius,
                double mass_fraction_cutoff, uint64_t seed, double* x, double* y,
                double* z, double* vx, double* vy, double* vz, double* m);
/* The same with the two directions of plummer.cpp:122 drawn in either order:
 * velocity_direction_first = 1 is a GCC build's right-to-left evaluation, the
 * order the reference's docs/performance/tree_accuracy.csv was sampled in. */
int orc_plummer_ordered(uint64_t count, double total_mass, double scale_radius,
                        double mass_fraction_cutoff, uint64_t seed, int velocity_direction_first,
                        double* x, double* y, double* z, double* vx, double* vy, double* vz,
                        double* m);

/* initial_conditions/uniform_sphere.cpp:32-61 (zero velocities, COM frame). */
int orc_uniform_sphere(uint64_t count, double total_mass, double radius, uint64_t seed,
                       double* x, double* y, double* z, double* vx, double* vy,
                       double* vz, double* m);

/* Two exponential discs with Hernquist bulges on a parabolic orbit: this
 * project's own generator for the two-disc collision workload (the reference's
 * disc generator uses a different vertical profile, bulge and orbit). */
typedef struct {
    uint64_t count;           /* total particles, split by mass ratio */
    double mass_ratio;        /* primary : secondary mass, e.g. 3 */
    double scale_length[2];   /* exponential scale lengths */
    double scale_height;      /* sech^2 vertical scale height */
    double truncation;        /* disc truncation in scale lengths */
    double bulge_fraction;    /* fraction of each galaxy's mass in its bulge */
    double bulge_scale;       /* Hernquist scale radius of the bulge */
    double dispersion;        /* Gaussian velocity dispersion, fraction of v_circ */
    double pericentre;        /* parabolic-orbit pericentre distance */
    double separation;        /* initial centre separation */
    double inclined_fraction; /* fraction of each disc's particles tilted */
    double inclination;       /* tilt of those particles, radians */
    double total_mass;        /* mass of both galaxies together */
    uint64_t seed;
} orc_disc_collision_params;
int orc_disc_collision(const orc_disc_collision_params* p, double* x, double* y, double* z,
                       double* vx, double* vy, double* vz, double* m);

/* ---- direct summation --------------------------------------------------- */

/* solvers/reference_kernel.cpp:31-104: FP64 per-target sum with Neumaier
 * compensation, explicit j == target skip, times G. `targets` may be NULL
 * (all targets, outputs in particle order) or a list of `target_count`
 * indices (outputs in list order). Threaded over targets. */
int orc_reference_acceleration(uint64_t n, const double* x, const double* y, const double* z,
                               const double* m, double softening_squared, double G,
                               const uint64_t* targets, uint64_t target_count, double* ax,
                               double* ay, double* az, unsigned threads);

/* solvers/direct_solver.cpp:20-80 over the kernel of direct_kernel_scalar.cpp:31-59
 * (kernel = 0) or direct_kernel_avx2.cpp:169-230 (kernel = 1): ranges [0,i) and
 * (i,N), a_i = G (before + after). Targets [target_begin, target_end) only,
 * outputs written at their particle index. Threaded with a work-stealing
 * executor restated from backend/work_stealing_executor.cpp:23-151. */
int orc_direct(uint64_t n, const double* x, const double* y, const double* z, const double* m,
               double softening_squared, double G, int kernel, unsigned threads,
               uint64_t target_begin, uint64_t target_end, double* ax, double* ay, double* az);

/* Single-precision build of the same solver (kernel = 1 uses 8 FP32 lanes). */
int orc_direct_f32(uint64_t n, const float* x, const float* y, const float* z, const float* m,
                   float softening_squared, float G, int kernel, unsigned threads,
                   uint64_t target_begin, uint64_t target_end, float* ax, float* ay, float* az);

/* ---- Morton ordering ---------------------------------------------------- */

/* solvers/morton.cpp:60-104 */
int orc_bounding_cube(uint64_t n, const double* x, const double* y, const double* z,
                      double origin[3], double* size);
/* solvers/morton.cpp:106-122 */
uint64_t orc_spread_bits(uint32_t value);
/* solvers/morton.cpp:124-126 and :166-178. */
uint64_t orc_morton_interleave(uint32_t x, uint32_t y, uint32_t z);
unsigned orc_morton_octant(uint64_t code, unsigned level);
/* solvers/morton.cpp:124-164: codes in particle order. */
int orc_morton_codes(uint64_t n, const double* x, const double* y, const double* z,
                     const double origin[3], double size, uint64_t* codes);
/* solvers/morton.cpp:180-197: cube, codes and the (code, index) total order. */
int orc_morton_order(uint64_t n, const double* x, const double* y, const double* z,
                     double origin[3], double* size, uint64_t* sorted_codes,
                     uint64_t* sorted_index);

/* ---- octree ------------------------------------------------------------- */

typedef struct orc_tree orc_tree;
/* solvers/octree.cpp:196-468 (serial emit), moments :75-185, radius :374-387.
 * Inputs are the sorted SoA and sorted codes; NULL on usage error. */
orc_tree* orc_tree_build(uint64_t n, const double* sx, const double* sy, const double* sz,
                         const double* sm, const uint64_t* sorted_codes,
                         const double origin[3], double size, double opening_angle,
                         uint64_t leaf_capacity, int quadrupole);
uint64_t orc_tree_node_count(const orc_tree* tree);
uint64_t orc_tree_leaf_count(const orc_tree* tree);
unsigned orc_tree_depth(const orc_tree* tree);
/* Any output pointer may be NULL. quad is 6 doubles per node (xx,xy,xz,yy,yz,zz). */
void orc_tree_nodes(const orc_tree* tree, double* mass, double* com_x, double* com_y,
                    double* com_z, double* acceptance_radius_squared, uint64_t* next,
                    uint64_t* first_particle, uint64_t* particle_count, double* quad);
void orc_tree_free(orc_tree* tree);

/* solvers/barnes_hut_solver.cpp:41-130 with tree_walk.cpp:20-83: order, gather,
 * build, walk each sorted target, scatter G a to particle order.
 * counts[0] = particle_particle, counts[1] = particle_cell. */
int orc_barnes_hut(uint64_t n, const double* x, const double* y, const double* z,
                   const double* m, double softening_squared, double G, double opening_angle,
                   uint64_t leaf_capacity, int quadrupole, int kernel, unsigned threads,
                   double* ax, double* ay, double* az, uint64_t counts[2]);

/* Interaction counts of the device walk's precision policy, for exact count
 * parity: the reference tree (FP64 moments and radii), then per target the
 * FP32 test r2 >= (float)acc_r2 with r2 = fma(dz, dz, fma(dy, dy, dx*dx)) on
 * FP32 positions relative to the root centre (origin + size/2). The decisions
 * follow tree_walk.cpp:42-81; only the arithmetic of the test is FP32.
 * counts[0] = particle_particle, counts[1] = particle_cell. */
int orc_barnes_hut_fp32_counts(uint64_t n, const double* x, const double* y, const double* z,
                               const double* m, double opening_angle, uint64_t leaf_capacity,
                               unsigned threads, uint64_t counts[2]);

/* ---- integrators and diagnostics ---------------------------------------- */

/* integrators/velocity_verlet.cpp:20-72 (integrator 0) and yoshida4.cpp:24-59
 * (integrator 1), with the direct solver (solver 0, scalar kernel, or 2, AVX2
 * kernel) or Barnes-Hut (solver 1) as the field. Accelerations on entry must
 * belong to the positions (integrator.hpp:51-53); orc_refresh computes them. */
typedef struct {
    int integrator;
    int solver;
    double softening_squared;
    double G;
    double opening_angle;
    uint64_t leaf_capacity;
    unsigned threads;
} orc_field;
int orc_refresh(uint64_t n, const double* x, const double* y, const double* z, const double* m,
                const orc_field* field, double* ax, double* ay, double* az);
int orc_steps(uint64_t n, double* x, double* y, double* z, double* vx, double* vy, double* vz,
              double* ax, double* ay, double* az, const double* m, const orc_field* field,
              double timestep, uint64_t steps);

/* core/diagnostics.cpp:85-200: out = {kinetic, potential, px, py, pz, lx, ly, lz}. */
int orc_diagnostics(uint64_t n, const double* x, const double* y, const double* z,
                    const double* vx, const double* vy, const double* vz, const double* m,
                    double softening_squared, double G, double out[8]);

#ifdef __cplusplus
}
#endif

#endif
