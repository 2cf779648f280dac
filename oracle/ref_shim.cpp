// C-ABI shim over the reference's own CPU library -- TEST INFRASTRUCTURE ONLY.
//
// Compiled by oracle/ref.mk against the reference headers and sources where
// they lie (never copied), into oracle/_ref/libref.so. Tests use it to pin
// the oracle restatement bit for bit; bench.py may time it as the CPU
// baseline ("kind": "reference"). The product never loads it.
//
// Softening is passed as a length here (the reference stores its square,
// core/softening.hpp:65), so callers square it the same way the reference does.
#include <cstdint>
#include <cstring>
#include <memory>
#include <span>
#include <sstream>
#include <string>
#include <vector>

#include "orrery/backend/work_stealing_executor.hpp"
#include "orrery/core/diagnostics.hpp"
#include "orrery/core/particle_data.hpp"
#include "orrery/core/random.hpp"
#include "orrery/core/softening.hpp"
#include "orrery/core/vec3_span.hpp"
#include "orrery/initial_conditions/plummer.hpp"
#include "orrery/initial_conditions/uniform_sphere.hpp"
#include "orrery/integrators/integrator.hpp"
#include "orrery/sim/checkpoint.hpp"
#include "orrery/sim/config_file.hpp"
#include "orrery/sim/trajectory.hpp"
#include "orrery/integrators/runge_kutta4.hpp"
#include "orrery/integrators/velocity_verlet.hpp"
#include "orrery/integrators/yoshida4.hpp"
#include "orrery/solvers/barnes_hut_solver.hpp"
#include "orrery/solvers/direct_kernel.hpp"
#include "orrery/solvers/direct_solver.hpp"
#include "orrery/solvers/morton.hpp"
#include "orrery/solvers/octree.hpp"
#include "orrery/solvers/reference_kernel.hpp"

using namespace orrery;

namespace {

core::Vec3Span<const double> view(std::uint64_t n, const double* x, const double* y,
                                  const double* z) {
    return {std::span<const double>(x, n), std::span<const double>(y, n),
            std::span<const double>(z, n)};
}
core::Vec3Span<double> view_mut(std::uint64_t n, double* x, double* y, double* z) {
    return {std::span<double>(x, n), std::span<double>(y, n), std::span<double>(z, n)};
}

void copy_data(const core::ParticleData& d, double* x, double* y, double* z, double* vx,
               double* vy, double* vz, double* m) {
    const auto n = d.size();
    const auto p = d.positions();
    const auto v = d.velocities();
    for (std::size_t i = 0; i < n; ++i) {
        x[i] = p.x[i];
        y[i] = p.y[i];
        z[i] = p.z[i];
        vx[i] = v.x[i];
        vy[i] = v.y[i];
        vz[i] = v.z[i];
        m[i] = d.masses()[i];
    }
}

core::ParticleData load(std::uint64_t n, const double* x, const double* y, const double* z,
                        const double* vx, const double* vy, const double* vz, const double* ax,
                        const double* ay, const double* az, const double* m) {
    core::ParticleData d(n);
    for (std::size_t i = 0; i < n; ++i) {
        d.positions().set(i, {x[i], y[i], z[i]});
        d.velocities().set(i, {vx[i], vy[i], vz[i]});
        d.accelerations().set(i, {ax[i], ay[i], az[i]});
        d.masses()[i] = m[i];
    }
    return d;
}

}  // namespace

extern "C" {

int ref_plummer(std::uint64_t count, double M, double a, double cutoff, std::uint64_t seed,
                double* x, double* y, double* z, double* vx, double* vy, double* vz, double* m) {
    core::RandomSource random{seed};
    initial_conditions::PlummerParameters p;
    p.count = count;
    p.total_mass = M;
    p.scale_radius = a;
    p.mass_fraction_cutoff = cutoff;
    const core::ParticleData d = initial_conditions::make_plummer_sphere(p, random);
    copy_data(d, x, y, z, vx, vy, vz, m);
    return 0;
}

int ref_uniform_sphere(std::uint64_t count, double M, double R, std::uint64_t seed, double* x,
                       double* y, double* z, double* vx, double* vy, double* vz, double* m) {
    core::RandomSource random{seed};
    initial_conditions::UniformSphereParameters p;
    p.count = count;
    p.total_mass = M;
    p.radius = R;
    const core::ParticleData d = initial_conditions::make_uniform_sphere(p, random);
    copy_data(d, x, y, z, vx, vy, vz, m);
    return 0;
}

int ref_reference_acceleration(std::uint64_t n, const double* x, const double* y,
                               const double* z, const double* m, double eps,
                               const std::uint64_t* targets, std::uint64_t tcount, double* ax,
                               double* ay, double* az) {
    const auto pos = view(n, x, y, z);
    const std::span<const double> mass(m, n);
    const std::uint64_t total = targets == nullptr ? n : tcount;
    for (std::uint64_t k = 0; k < total; ++k) {
        const std::uint64_t t = targets == nullptr ? k : targets[k];
        const auto r = solvers::reference_acceleration(pos, mass, t, core::Softening{eps});
        ax[k] = r.x;
        ay[k] = r.y;
        az[k] = r.z;
    }
    return 0;
}

int ref_direct(std::uint64_t n, const double* x, const double* y, const double* z,
               const double* m, double eps, int kernel, unsigned threads, double* ax, double* ay,
               double* az) {
    std::unique_ptr<backend::WorkStealingExecutor> ex;
    std::unique_ptr<solvers::DirectSolver> solver;
    if (threads == 1) {
        solver = std::make_unique<solvers::DirectSolver>(core::Softening{eps});
    } else {
        ex = std::make_unique<backend::WorkStealingExecutor>(threads);
        solver = std::make_unique<solvers::DirectSolver>(core::Softening{eps}, *ex);
    }
    solver->select_kernel(kernel == 1 ? solvers::KernelKind::kAvx2 : solvers::KernelKind::kScalar);
    solver->evaluate(view(n, x, y, z), std::span<const double>(m, n), view_mut(n, ax, ay, az));
    return 0;
}

// The body of DirectSolver::evaluate (solvers/direct_solver.cpp:39-58) over a
// sub-range of targets, with the reference's own kernels and work-stealing
// executor: a bounded sample of a large direct evaluation for CPU timing.
int ref_direct_targets(std::uint64_t n, const double* x, const double* y, const double* z,
                       const double* m, double eps, int kernel, unsigned threads,
                       std::uint64_t tb, std::uint64_t te, double* ax, double* ay, double* az) {
    const auto pos = view(n, x, y, z);
    const std::span<const double> mass(m, n);
    const core::Softening softening{eps};
    const solvers::AccumulateRange accumulate = solvers::accumulate_range_for(
        kernel == 1 ? solvers::KernelKind::kAvx2 : solvers::KernelKind::kScalar);
    backend::WorkStealingExecutor ex(threads);
    ex.run(te - tb, [&](core::Index b, core::Index e) {
        for (core::Index k = b; k < e; ++k) {
            const core::Index i = tb + k;
            const core::Vec3 t = pos.get(i);
            const core::Vec3 before = accumulate(pos, mass, t, 0, i, softening);
            const core::Vec3 after = accumulate(pos, mass, t, i + 1, n, softening);
            const core::Vec3 a = before + after;
            ax[i] = a.x;
            ay[i] = a.y;
            az[i] = a.z;
        }
    });
    return 0;
}

int ref_morton_order(std::uint64_t n, const double* x, const double* y, const double* z,
                     double origin[3], double* size, std::uint64_t* codes,
                     std::uint64_t* index) {
    solvers::MortonOrdering ordering;
    ordering.build(view(n, x, y, z), nullptr);
    const auto keys = ordering.keys();
    for (std::size_t i = 0; i < n; ++i) {
        codes[i] = keys[i].code;
        index[i] = keys[i].index;
    }
    origin[0] = ordering.cube().origin.x;
    origin[1] = ordering.cube().origin.y;
    origin[2] = ordering.cube().origin.z;
    *size = ordering.cube().size;
    return 0;
}

// Builds the tree over the sorted arrays exactly as BarnesHutSolver does
// (barnes_hut_solver.cpp:64-80) and copies the nodes out. Returns the node
// count; arrays must hold `capacity` nodes (nothing is written past it).
std::int64_t ref_tree(std::uint64_t n, const double* x, const double* y, const double* z,
                      const double* m, double theta, std::uint64_t cap, int quadrupole,
                      unsigned threads, std::uint64_t capacity, double* mass, double* cx,
                      double* cy, double* cz, double* acc, std::uint64_t* next,
                      std::uint64_t* first, std::uint64_t* count, double* quad,
                      std::uint64_t* leaves_depth) {
    std::unique_ptr<backend::WorkStealingExecutor> ex;
    if (threads != 1) ex = std::make_unique<backend::WorkStealingExecutor>(threads);
    solvers::MortonOrdering ordering;
    ordering.build(view(n, x, y, z), ex.get());
    const auto keys = ordering.keys();
    std::vector<double> sx(n), sy(n), sz(n), sm(n);
    for (std::size_t i = 0; i < n; ++i) {
        sx[i] = x[keys[i].index];
        sy[i] = y[keys[i].index];
        sz[i] = z[keys[i].index];
        sm[i] = m[keys[i].index];
    }
    solvers::TreeParameters p;
    p.opening_angle = theta;
    p.leaf_capacity = cap;
    p.quadrupole = quadrupole != 0;
    solvers::Octree tree;
    tree.build(view(n, sx.data(), sy.data(), sz.data()), std::span<const double>(sm), keys,
               ordering.cube(), p, ex.get());
    const auto nodes = tree.nodes();
    if (nodes.size() <= capacity) {
        for (std::size_t i = 0; i < nodes.size(); ++i) {
            mass[i] = nodes[i].mass;
            cx[i] = nodes[i].centre_of_mass.x;
            cy[i] = nodes[i].centre_of_mass.y;
            cz[i] = nodes[i].centre_of_mass.z;
            acc[i] = nodes[i].acceptance_radius_squared;
            next[i] = nodes[i].next;
            first[i] = nodes[i].first_particle;
            count[i] = nodes[i].particle_count;
            if (quad != nullptr && p.quadrupole) {
                const auto& q = tree.quadrupoles()[i];
                double* o = quad + 6 * i;
                o[0] = q.xx; o[1] = q.xy; o[2] = q.xz; o[3] = q.yy; o[4] = q.yz; o[5] = q.zz;
            }
        }
    }
    leaves_depth[0] = tree.leaf_count();
    leaves_depth[1] = tree.depth();
    return static_cast<std::int64_t>(nodes.size());
}

int ref_barnes_hut(std::uint64_t n, const double* x, const double* y, const double* z,
                   const double* m, double eps, double theta, std::uint64_t cap, int quadrupole,
                   int kernel, unsigned threads, double* ax, double* ay, double* az,
                   std::uint64_t counts[2]) {
    solvers::TreeParameters p;
    p.opening_angle = theta;
    p.leaf_capacity = cap;
    p.quadrupole = quadrupole != 0;
    std::unique_ptr<backend::WorkStealingExecutor> ex;
    std::unique_ptr<solvers::BarnesHutSolver> solver;
    if (threads == 1) {
        solver = std::make_unique<solvers::BarnesHutSolver>(p, core::Softening{eps});
    } else {
        ex = std::make_unique<backend::WorkStealingExecutor>(threads);
        solver = std::make_unique<solvers::BarnesHutSolver>(p, core::Softening{eps}, *ex);
    }
    solver->select_kernel(kernel == 1 ? solvers::KernelKind::kAvx2 : solvers::KernelKind::kScalar);
    solver->evaluate(view(n, x, y, z), std::span<const double>(m, n), view_mut(n, ax, ay, az));
    counts[0] = solver->interaction_count().particle_particle;
    counts[1] = solver->interaction_count().particle_cell;
    return 0;
}

// Runs `steps` steps of VV (integrator 0), Yoshida-4 (1) or RK4 (2) with the
// direct solver (solver 0 scalar, 2 AVX2) or Barnes-Hut (1). `threads` > 1
// puts the solver on a work-stealing executor of that many workers, the way
// sim/assembly.cpp:94-107 and :233-243 assemble a threaded CPU run; 1 runs it
// on the calling thread.
int ref_steps_threaded(std::uint64_t n, double* x, double* y, double* z, double* vx,
                       double* vy, double* vz, double* ax, double* ay, double* az,
                       const double* m, int integrator, int solver_kind, double eps,
                       double theta, std::uint64_t cap, double dt, std::uint64_t steps,
                       unsigned threads) {
    core::ParticleData d = load(n, x, y, z, vx, vy, vz, ax, ay, az, m);
    std::unique_ptr<backend::WorkStealingExecutor> ex;
    if (threads > 1) ex = std::make_unique<backend::WorkStealingExecutor>(threads);
    std::unique_ptr<solvers::ForceSolver> solver;
    if (solver_kind == 1) {
        solvers::TreeParameters p;
        p.opening_angle = theta;
        p.leaf_capacity = cap;
        auto bh = ex ? std::make_unique<solvers::BarnesHutSolver>(p, core::Softening{eps}, *ex)
                     : std::make_unique<solvers::BarnesHutSolver>(p, core::Softening{eps});
        bh->select_kernel(solvers::KernelKind::kAvx2);
        solver = std::move(bh);
    } else {
        auto ds = ex ? std::make_unique<solvers::DirectSolver>(core::Softening{eps}, *ex)
                     : std::make_unique<solvers::DirectSolver>(core::Softening{eps});
        ds->select_kernel(solver_kind == 2 ? solvers::KernelKind::kAvx2
                                           : solvers::KernelKind::kScalar);
        solver = std::move(ds);
    }
    std::unique_ptr<integrators::Integrator> step;
    if (integrator == 2) {
        step = std::make_unique<integrators::RungeKutta4>();
    } else if (integrator == 1) {
        step = std::make_unique<integrators::Yoshida4>();
    } else {
        step = std::make_unique<integrators::VelocityVerlet>();
    }
    for (std::uint64_t s = 0; s < steps; ++s) step->step(d, dt, *solver);
    const auto p = d.positions();
    const auto v = d.velocities();
    const auto a = d.accelerations();
    for (std::size_t i = 0; i < n; ++i) {
        x[i] = p.x[i]; y[i] = p.y[i]; z[i] = p.z[i];
        vx[i] = v.x[i]; vy[i] = v.y[i]; vz[i] = v.z[i];
        ax[i] = a.x[i]; ay[i] = a.y[i]; az[i] = a.z[i];
    }
    return 0;
}

// Single-threaded form of ref_steps_threaded.
int ref_steps(std::uint64_t n, double* x, double* y, double* z, double* vx, double* vy,
              double* vz, double* ax, double* ay, double* az, const double* m, int integrator,
              int solver_kind, double eps, double theta, std::uint64_t cap, double dt,
              std::uint64_t steps) {
    return ref_steps_threaded(n, x, y, z, vx, vy, vz, ax, ay, az, m, integrator, solver_kind,
                              eps, theta, cap, dt, steps, 1u);
}

int ref_diagnostics(std::uint64_t n, const double* x, const double* y, const double* z,
                    const double* vx, const double* vy, const double* vz, const double* m,
                    double eps, double out[8]) {
    std::vector<double> zero(n, 0.0);
    const core::ParticleData d =
        load(n, x, y, z, vx, vy, vz, zero.data(), zero.data(), zero.data(), m);
    const core::Diagnostics g = core::measure_diagnostics(d, core::Softening{eps});
    out[0] = g.kinetic_energy;
    out[1] = g.potential_energy;
    out[2] = g.linear_momentum.x;
    out[3] = g.linear_momentum.y;
    out[4] = g.linear_momentum.z;
    out[5] = g.angular_momentum.x;
    out[6] = g.angular_momentum.y;
    out[7] = g.angular_momentum.z;
    return 0;
}


// sim/checkpoint.cpp through the reference's own reader: step, count, the ten
// arrays (m, x, y, z, vx, vy, vz, ax, ay, az, each `capacity` long) and the
// configuration as the reference writes it back. 0 on success, 1 on refusal.
int ref_checkpoint_read(const char* path, std::uint64_t* step, std::uint64_t* count,
                        std::uint64_t capacity, double* const* arrays, char* text,
                        std::uint64_t text_capacity) {
    try {
        const sim::Checkpoint c = sim::read_checkpoint(path);
        *step = c.step;
        *count = c.particles.size();
        if (*count > capacity) return 2;
        const auto& d = c.particles;
        const std::span<const double> src[10] = {
            d.masses(), d.positions().x, d.positions().y, d.positions().z,
            d.velocities().x, d.velocities().y, d.velocities().z,
            d.accelerations().x, d.accelerations().y, d.accelerations().z};
        for (int k = 0; k < 10; ++k) {
            for (std::size_t i = 0; i < *count; ++i) arrays[k][i] = src[k][i];
        }
        std::ostringstream out;
        sim::write_configuration(out, c.configuration);
        const std::string t = out.str();
        const std::size_t n = std::min<std::size_t>(t.size(), text_capacity - 1);
        std::memcpy(text, t.data(), n);
        text[n] = '\0';
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

// A checkpoint written by the reference: configuration parsed from `text`.
int ref_checkpoint_write(const char* path, std::uint64_t step, std::uint64_t n,
                         const double* const* arrays, const char* text) {
    try {
        sim::Checkpoint c;
        std::istringstream in(text);
        c.configuration = sim::parse_configuration(in, "shim");
        c.step = step;
        c.particles.resize(n);
        auto& d = c.particles;
        const std::span<double> dst[10] = {
            d.masses(), d.positions().x, d.positions().y, d.positions().z,
            d.velocities().x, d.velocities().y, d.velocities().z,
            d.accelerations().x, d.accelerations().y, d.accelerations().z};
        for (int k = 0; k < 10; ++k) {
            for (std::size_t i = 0; i < n; ++i) dst[k][i] = arrays[k][i];
        }
        sim::write_checkpoint(path, c);
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

// Frames of a trajectory through the reference's TrajectoryReader: the header
// fields, then up to max_frames frames of step, time and 6 x n values
// (x y z, then vx vy vz when present). Returns the frame count, or -1.
long ref_trajectory_read(const char* path, std::uint64_t* count, double* timestep,
                         int* has_velocities, double* masses, std::uint64_t capacity,
                         std::uint64_t max_frames, std::uint64_t* steps, double* times,
                         double* values) {
    try {
        sim::TrajectoryReader r(path);
        *count = r.info().particle_count;
        *timestep = r.info().timestep;
        *has_velocities = r.info().has_velocities ? 1 : 0;
        const std::size_t n = *count;
        if (n > capacity) return -1;
        for (std::size_t i = 0; i < n; ++i) masses[i] = r.masses()[i];
        sim::TrajectoryFrame f;
        long frames = 0;
        while (static_cast<std::uint64_t>(frames) < max_frames && r.read_frame(f)) {
            steps[frames] = f.step;
            times[frames] = f.time;
            double* v = values + static_cast<std::size_t>(frames) * 6 * n;
            const auto p = f.positions.view();
            for (std::size_t i = 0; i < n; ++i) {
                v[i] = p.x[i];
                v[n + i] = p.y[i];
                v[2 * n + i] = p.z[i];
            }
            if (*has_velocities != 0) {
                const auto u = f.velocities.view();
                for (std::size_t i = 0; i < n; ++i) {
                    v[3 * n + i] = u.x[i];
                    v[4 * n + i] = u.y[i];
                    v[5 * n + i] = u.z[i];
                }
            }
            ++frames;
        }
        return frames;
    } catch (const std::exception&) {
        return -1;
    }
}
}  // extern "C"
