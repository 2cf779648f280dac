// C++ tests of the ForceSolver drop-ins (HipDirectSolver, HipTreeSolver).
//
// Mirrors the reference's tests/solvers/direct_solver_test.cpp (closed forms,
// write-not-accumulate, self term, empty input) and cuda_direct_solver_test.cpp
// (try_create returns null without a device; agreement with an FP64 sum).
// Usage: test_host_solvers [--expect-no-device]. Exit status 0 on success.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <span>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "orrery_hip/hip_assembly.hpp"
#include "orrery_hip/hip_solvers.hpp"

using orrery::core::Real;
using orrery::core::Softening;
using orrery::core::Vec3Span;
using orrery::solvers::HipDirectSolver;
using orrery::solvers::HipSolverKind;
using orrery::solvers::HipSolverSettings;
using orrery::solvers::HipTreeSolver;
using orrery::solvers::TreeParameters;

namespace {

int failures = 0;

#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::fprintf(stderr, "%s:%d: CHECK failed: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                          \
        }                                                                        \
    } while (0)

struct System {
    std::vector<Real> x, y, z, m, ax, ay, az;
    explicit System(std::size_t n) : x(n), y(n), z(n), m(n), ax(n), ay(n), az(n) {}
    Vec3Span<const Real> positions() const { return {std::span<const Real>(x), std::span<const Real>(y), std::span<const Real>(z)}; }
    Vec3Span<Real> accelerations() { return {std::span<Real>(ax), std::span<Real>(ay), std::span<Real>(az)}; }
    std::span<const Real> masses() const { return m; }
};

// Plain FP64 direct sum for the agreement check (test-side reference).
void fp64_direct(const System& s, double eps2, std::vector<double>& rx, std::vector<double>& ry,
                 std::vector<double>& rz) {
    const std::size_t n = s.m.size();
    rx.assign(n, 0.0);
    ry.assign(n, 0.0);
    rz.assign(n, 0.0);
    for (std::size_t i = 0; i < n; ++i) {
        long double a[3] = {0, 0, 0};
        for (std::size_t j = 0; j < n; ++j) {
            if (j == i) continue;
            const double dx = s.x[j] - s.x[i], dy = s.y[j] - s.y[i], dz = s.z[j] - s.z[i];
            const double inv = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz + eps2);
            const double f = s.m[j] * inv * inv * inv;
            a[0] += f * dx;
            a[1] += f * dy;
            a[2] += f * dz;
        }
        rx[i] = static_cast<double>(a[0]);
        ry[i] = static_cast<double>(a[1]);
        rz[i] = static_cast<double>(a[2]);
    }
}

double rms_relative(const System& s, const std::vector<double>& rx, const std::vector<double>& ry,
                    const std::vector<double>& rz) {
    double total = 0;
    for (std::size_t i = 0; i < s.m.size(); ++i) {
        const double dx = s.ax[i] - rx[i], dy = s.ay[i] - ry[i], dz = s.az[i] - rz[i];
        const double mag = std::sqrt(rx[i] * rx[i] + ry[i] * ry[i] + rz[i] * rz[i]);
        if (mag > 0) total += (dx * dx + dy * dy + dz * dz) / (mag * mag);
    }
    return std::sqrt(total / static_cast<double>(s.m.size()));
}

System random_cluster(std::size_t n, unsigned seed) {
    System s(n);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> g(0.0, 1.0);
    for (std::size_t i = 0; i < n; ++i) {
        s.x[i] = g(rng);
        s.y[i] = g(rng);
        s.z[i] = g(rng);
        s.m[i] = 1.0 / static_cast<double>(n);
    }
    return s;
}

void direct_cases() {
    auto solver = HipDirectSolver::try_create(Softening{});
    CHECK(solver != nullptr);
    if (!solver) return;
    CHECK(solver->name() == "hip-direct");
    // direct_solver_test.cpp:93-114 and :182-199 (written, not accumulated).
    System s(2);
    s.x = {-1, 1};
    s.m = {2, 3};
    s.ax = {7, -9};
    s.ay = {7, 9};
    solver->evaluate(s.positions(), s.masses(), s.accelerations());
    CHECK(s.ax[0] == 0.75 && s.ax[1] == -0.5 && s.ay[0] == 0 && s.az[1] == 0);
    // Self term (direct_solver_test.cpp:201-219).
    System one(1);
    one.x = {2};
    one.y = {-3};
    one.z = {5};
    one.m = {4};
    solver->evaluate(one.positions(), one.masses(), one.accelerations());
    CHECK(one.ax[0] == 0 && one.ay[0] == 0 && one.az[0] == 0);
    // Empty input and counts.
    System none(0);
    solver->evaluate(none.positions(), none.masses(), none.accelerations());
    CHECK(solver->interaction_count().evaluations == 3);
    CHECK(solver->interaction_count().particle_particle == 2);
    solver->reset_interaction_count();
    CHECK(solver->interaction_count().evaluations == 0);
    // Plummer softening factor (direct_solver_test.cpp:142-178).
    auto soft = HipDirectSolver::try_create(Softening{3});
    System pair(2);
    pair.x = {0, 4};
    pair.m = {1, 5};
    soft->evaluate(pair.positions(), pair.masses(), pair.accelerations());
    const double expected = 5.0 * 4.0 / std::pow(25.0, 1.5);
    CHECK(std::abs(pair.ax[0] - expected) <= 4e-7 * expected);
    CHECK(soft->softening().squared() == 9.0);
    // Agreement with FP64 on a random cluster.
    System c = random_cluster(3000, 7);
    auto s2 = HipDirectSolver::try_create(Softening{0.05});
    s2->set_timing(true);
    s2->evaluate(c.positions(), c.masses(), c.accelerations());
    std::vector<double> rx, ry, rz;
    fp64_direct(c, 0.05 * 0.05, rx, ry, rz);
    const double rms = rms_relative(c, rx, ry, rz);
    std::printf("hip-direct N=3000 rms relative error %.3e\n", rms);
    CHECK(rms < 1e-6);
    CHECK(s2->timings().kernel > 0);
    // Ordered pairs, N(N-1).
    CHECK(s2->interaction_count().particle_particle == 3000ull * 2999ull);
    // Timing off (the default): later evaluations of one size replay the
    // captured ordering graph and give the same bits as the timed, eager one.
    s2->set_timing(false);
    s2->reset_timings();
    const std::vector<double> eager_x = c.ax, eager_y = c.ay, eager_z = c.az;
    for (int k = 0; k < 4; ++k) {
        s2->evaluate(c.positions(), c.masses(), c.accelerations());
        CHECK(c.ax == eager_x && c.ay == eager_y && c.az == eager_z);
    }
    CHECK(s2->timings().graph_launches >= 2);
    CHECK(s2->timings().kernel == 0);
    // Mismatched spans are a usage error.
    System bad(3);
    bad.m.resize(2);
    bool threw = false;
    try {
        s2->evaluate(bad.positions(), bad.masses(), bad.accelerations());
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);
}

void tree_cases() {
    TreeParameters p;
    p.opening_angle = 0.5;
    p.leaf_capacity = 16;
    auto solver = HipTreeSolver::try_create(p, Softening{0.05});
    CHECK(solver != nullptr);
    if (!solver) return;
    CHECK(solver->name() == "hip-tree");
    System c = random_cluster(5000, 11);
    solver->evaluate(c.positions(), c.masses(), c.accelerations());
    std::vector<double> rx, ry, rz;
    fp64_direct(c, 0.05 * 0.05, rx, ry, rz);
    const double rms = rms_relative(c, rx, ry, rz);
    std::printf("hip-tree N=5000 theta 0.5 rms relative error %.3e, nodes %llu, depth %u\n", rms,
                static_cast<unsigned long long>(solver->tree().node_count), solver->tree().depth);
    CHECK(rms < 1e-2);
    CHECK(solver->interaction_count().particle_particle > 0);
    CHECK(solver->interaction_count().particle_cell > 0);
    CHECK(solver->tree().leaf_count > 0 && solver->tree().leaf_count <= solver->tree().node_count);
    // Untimed by default: the second evaluation of a size is captured and the
    // later ones replay the graph, bit for bit the eager result; with timing
    // on the evaluation runs eagerly, again with the same bits.
    const std::vector<double> first_x = c.ax, first_y = c.ay, first_z = c.az;
    solver->reset_timings();
    for (int k = 0; k < 4; ++k) {
        solver->evaluate(c.positions(), c.masses(), c.accelerations());
        CHECK(c.ax == first_x && c.ay == first_y && c.az == first_z);
    }
    CHECK(solver->timings().graph_launches >= 2);
    CHECK(solver->timings().traversal == 0);
    solver->set_timing(true);
    solver->evaluate(c.positions(), c.masses(), c.accelerations());
    CHECK(c.ax == first_x && c.ay == first_y && c.az == first_z);
    CHECK(solver->timings().traversal > 0);
    solver->set_timing(false);
    // theta = 0: the direct sum, N(N-1) pairs, no cells (barnes_hut_test.cpp:111-152).
    p.opening_angle = 0;
    auto exact = HipTreeSolver::try_create(p, Softening{0.05});
    exact->evaluate(c.positions(), c.masses(), c.accelerations());
    CHECK(exact->interaction_count().particle_particle == 5000ull * 4999ull);
    CHECK(exact->interaction_count().particle_cell == 0);
    CHECK(rms_relative(c, rx, ry, rz) < 1e-6);
    // Parameters are corrected as octree.cpp:189-194 does.
    TreeParameters wild;
    wild.opening_angle = 3;
    wild.leaf_capacity = 0;
    auto corrected = HipTreeSolver::try_create(wild);
    CHECK(corrected->parameters().opening_angle == 1 && corrected->parameters().leaf_capacity == 1);
}

// `[solver] kind` spellings and the assembly step (assembly.cpp:116-129, 246-272).
void assembly_cases(bool device) {
    using orrery::solvers::cpu_equivalent;
    using orrery::solvers::parse_hip_solver_kind;
    CHECK(parse_hip_solver_kind("hip-direct") == HipSolverKind::kHipDirect);
    CHECK(parse_hip_solver_kind("hip-tree") == HipSolverKind::kHipTree);
    CHECK(!parse_hip_solver_kind("HIP-tree").has_value());
    CHECK(!parse_hip_solver_kind("hip-tree ").has_value());
    CHECK(!parse_hip_solver_kind("sycl-tree").has_value());
    CHECK(orrery::solvers::to_string(HipSolverKind::kHipTree) == "hip-tree");
    CHECK(cpu_equivalent(HipSolverKind::kHipTree) == "barnes-hut");
    CHECK(cpu_equivalent(HipSolverKind::kHipDirect) == "direct");
    CHECK(orrery::solvers::is_tree_solver(HipSolverKind::kHipTree));
    CHECK(!orrery::solvers::is_tree_solver(HipSolverKind::kHipDirect));

    HipSolverSettings tree;
    tree.kind = HipSolverKind::kHipTree;
    tree.softening = 0.05;
    tree.opening_angle = 1.7;  // corrected to 1 (octree.cpp:189-194)
    tree.leaf_capacity = 0;    // corrected to 1
    tree.quadrupole = true;
    HipSolverSettings direct;
    direct.kind = HipSolverKind::kHipDirect;
    direct.softening = 0.1;
    std::ostringstream report;
    if (device) {
        auto t = orrery::solvers::make_hip_solver(tree, report);
        auto d = orrery::solvers::make_hip_solver(direct, report);
        CHECK(t != nullptr && d != nullptr && report.str().empty());
        if (t == nullptr || d == nullptr) return;
        CHECK(t->name() == "hip-tree" && d->name() == "hip-direct");
        CHECK(t->softening().squared() == 0.05 * 0.05 && d->softening().squared() == 0.1 * 0.1);
        const auto* ht = dynamic_cast<const HipTreeSolver*>(t.get());
        CHECK(ht != nullptr);
        if (ht != nullptr) {
            CHECK(ht->parameters().opening_angle == 1.0);
            CHECK(ht->parameters().leaf_capacity == 1);
            CHECK(ht->parameters().quadrupole);
        }
        return;
    }
    CHECK(orrery::solvers::try_make_hip_solver(tree) == nullptr);
    // Fallback allowed: a null and the reference's note, naming the CPU solver.
    CHECK(orrery::solvers::make_hip_solver(tree, report) == nullptr);
    CHECK(report.str() ==
          "note: hip-tree was asked for but no usable device was found, so this run uses "
          "barnes-hut on the CPU\n");
    // Fallback refused: the run does not start.
    direct.allow_cpu_fallback = false;
    bool refused = false;
    try {
        (void)orrery::solvers::make_hip_solver(direct, report);
    } catch (const std::runtime_error& e) {
        refused = std::string(e.what()) ==
                  "solver.kind is hip-direct and no usable device was found. Set "
                  "solver.allow_cpu_fallback = true to run on the CPU instead";
    }
    CHECK(refused);
}

}  // namespace

int main(int argc, char** argv) {
    const bool expect_no_device = argc > 1 && std::strcmp(argv[1], "--expect-no-device") == 0;
    if (expect_no_device) {
        // cuda_direct_solver_test.cpp:132-136: no device -> null, the caller falls back.
        CHECK(HipDirectSolver::try_create(Softening{0.1}) == nullptr);
        CHECK(HipTreeSolver::try_create() == nullptr);
        assembly_cases(false);
    } else {
        direct_cases();
        tree_cases();
        assembly_cases(true);
    }
    if (failures == 0) std::printf("all host-solver checks passed\n");
    return failures == 0 ? 0 : 1;
}
