// Drives every entry point of the oracle (oracle/oracle.h) on small inputs and
// edge cases, for a build under AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/test_host_sanitizers.py). Values are not checked here (the oracle's
// own tests do that); the point is that no call reads or writes out of bounds,
// leaks, or hits undefined behaviour, threaded or not.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "oracle.h"

namespace {

int failures = 0;

#define CHECK(cond)                                                          \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::printf("%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                      \
        }                                                                    \
    } while (0)

struct State {
    explicit State(std::size_t n) : x(n), y(n), z(n), vx(n), vy(n), vz(n), m(n), ax(n), ay(n), az(n) {}
    std::vector<double> x, y, z, vx, vy, vz, m, ax, ay, az;
    std::size_t n() const { return x.size(); }
};

void solvers_on(State& s, unsigned threads) {
    const std::uint64_t n = s.n();
    const double eps2 = 1e-4;
    CHECK(orc_reference_acceleration(n, s.x.data(), s.y.data(), s.z.data(), s.m.data(), eps2, 1.0,
                                     nullptr, 0, s.ax.data(), s.ay.data(), s.az.data(), threads) == 0);
    if (n > 2) {
        const std::uint64_t targets[3] = {0, n / 2, n - 1};
        double tx[3], ty[3], tz[3];
        CHECK(orc_reference_acceleration(n, s.x.data(), s.y.data(), s.z.data(), s.m.data(), eps2, 1.0,
                                         targets, 3, tx, ty, tz, threads) == 0);
    }
    for (int kernel = 0; kernel < 2; ++kernel) {
        CHECK(orc_direct(n, s.x.data(), s.y.data(), s.z.data(), s.m.data(), eps2, 1.0, kernel, threads,
                         0, n, s.ax.data(), s.ay.data(), s.az.data()) == 0);
        CHECK(orc_direct(n, s.x.data(), s.y.data(), s.z.data(), s.m.data(), eps2, 1.0, kernel, threads,
                         n / 3, n / 2, s.ax.data(), s.ay.data(), s.az.data()) == 0);
        std::vector<float> fx(s.x.begin(), s.x.end()), fy(s.y.begin(), s.y.end()),
            fz(s.z.begin(), s.z.end()), fm(s.m.begin(), s.m.end()), gx(n), gy(n), gz(n);
        CHECK(orc_direct_f32(n, fx.data(), fy.data(), fz.data(), fm.data(), 1e-4f, 1.f, kernel, threads,
                             0, n, gx.data(), gy.data(), gz.data()) == 0);
    }
    double origin[3], size = 0;
    CHECK(orc_bounding_cube(n, s.x.data(), s.y.data(), s.z.data(), origin, &size) == 0);
    std::vector<std::uint64_t> codes(n), sorted(n), index(n);
    CHECK(orc_morton_codes(n, s.x.data(), s.y.data(), s.z.data(), origin, size, codes.data()) == 0);
    CHECK(orc_morton_order(n, s.x.data(), s.y.data(), s.z.data(), origin, &size, sorted.data(),
                           index.data()) == 0);
    for (int quad = 0; quad < 2; ++quad) {
        std::vector<double> sx(n), sy(n), sz(n), sm(n);
        for (std::uint64_t i = 0; i < n; ++i) {
            sx[i] = s.x[index[i]];
            sy[i] = s.y[index[i]];
            sz[i] = s.z[index[i]];
            sm[i] = s.m[index[i]];
        }
        orc_tree* tree = orc_tree_build(n, sx.data(), sy.data(), sz.data(), sm.data(), sorted.data(),
                                        origin, size, 0.5, 4, quad);
        if (n > 0) {
            CHECK(tree != nullptr);
            const std::uint64_t M = orc_tree_node_count(tree);
            (void)orc_tree_leaf_count(tree);
            (void)orc_tree_depth(tree);
            std::vector<double> mass(M), cx(M), cy(M), cz(M), acc(M), q(6 * M);
            std::vector<std::uint64_t> next(M), first(M), count(M);
            orc_tree_nodes(tree, mass.data(), cx.data(), cy.data(), cz.data(), acc.data(), next.data(),
                           first.data(), count.data(), quad ? q.data() : nullptr);
            orc_tree_nodes(tree, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                           nullptr);
        }
        orc_tree_free(tree);
        for (int kernel = 0; kernel < 2; ++kernel) {
            std::uint64_t counts[2] = {0, 0};
            CHECK(orc_barnes_hut(n, s.x.data(), s.y.data(), s.z.data(), s.m.data(), eps2, 1.0, 0.5, 4,
                                 quad, kernel, threads, s.ax.data(), s.ay.data(), s.az.data(),
                                 counts) == 0);
        }
    }
    std::uint64_t counts[2] = {0, 0};
    CHECK(orc_barnes_hut_fp32_counts(n, s.x.data(), s.y.data(), s.z.data(), s.m.data(), 0.5, 4, threads,
                                     counts) == 0);
    for (int integrator = 0; integrator < 2; ++integrator) {
        for (int solver = 0; solver < 3; ++solver) {
            const orc_field field{integrator, solver, eps2, 1.0, 0.5, 4, threads};
            CHECK(orc_refresh(n, s.x.data(), s.y.data(), s.z.data(), s.m.data(), &field, s.ax.data(),
                              s.ay.data(), s.az.data()) == 0);
            CHECK(orc_steps(n, s.x.data(), s.y.data(), s.z.data(), s.vx.data(), s.vy.data(), s.vz.data(),
                            s.ax.data(), s.ay.data(), s.az.data(), s.m.data(), &field, 1e-3, 2) == 0);
        }
    }
    double out[8];
    CHECK(orc_diagnostics(n, s.x.data(), s.y.data(), s.z.data(), s.vx.data(), s.vy.data(), s.vz.data(),
                          s.m.data(), eps2, 1.0, out) == 0);
}

}  // namespace

int main() {
    double u[64];
    CHECK(orc_random_uniform(7, 64, u) == 0);
    (void)orc_spread_bits(0x1fffffu);
    (void)orc_morton_interleave(1u, 2u, 3u);
    (void)orc_morton_octant(~0ull >> 1, 21);
    (void)orc_morton_octant(0, 0);

    State plummer(300);
    CHECK(orc_plummer(300, 1.0, 1.0, 0.999, 1, plummer.x.data(), plummer.y.data(), plummer.z.data(),
                      plummer.vx.data(), plummer.vy.data(), plummer.vz.data(), plummer.m.data()) == 0);
    State sphere(200);
    CHECK(orc_uniform_sphere(200, 1.0, 1.0, 2, sphere.x.data(), sphere.y.data(), sphere.z.data(),
                             sphere.vx.data(), sphere.vy.data(), sphere.vz.data(), sphere.m.data()) == 0);
    State discs(400);
    orc_disc_collision_params p{};
    p.count = 400;
    p.mass_ratio = 3;
    p.scale_length[0] = 1.0;
    p.scale_length[1] = 0.7;
    p.scale_height = 0.1;
    p.truncation = 5;
    p.bulge_fraction = 0.2;
    p.bulge_scale = 0.2;
    p.dispersion = 0.1;
    p.pericentre = 2;
    p.separation = 20;
    p.inclined_fraction = 0.3;
    p.inclination = 0.5235987755982988;
    p.total_mass = 1;
    p.seed = 3;
    CHECK(orc_disc_collision(&p, discs.x.data(), discs.y.data(), discs.z.data(), discs.vx.data(),
                             discs.vy.data(), discs.vz.data(), discs.m.data()) == 0);

    for (unsigned threads : {1u, 4u}) {
        solvers_on(plummer, threads);
        solvers_on(sphere, threads);
        solvers_on(discs, threads);
    }
    // Edge cases: one particle, coincident particles, and a massless set.
    State one(1);
    one.m[0] = 1;
    solvers_on(one, 2);
    State same(40);
    for (std::size_t i = 0; i < same.n(); ++i) same.m[i] = 1.0 / 40;
    solvers_on(same, 4);
    State massless(50);
    for (std::size_t i = 0; i < massless.n(); ++i) massless.x[i] = static_cast<double>(i) / 50;
    solvers_on(massless, 4);

    if (failures == 0) std::printf("oracle sanitizer run passed\n");
    return failures == 0 ? 0 : 1;
}
