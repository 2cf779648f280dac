// CPU oracle for the MI355X force engine -- TEST INFRASTRUCTURE ONLY.
//
// Loaded by tests/, the smoke check and bench.py's CPU-baseline leg, and by
// nothing in the product. Restates the reference CPU path in this project's
// own code; every function cites the reference file:line it follows.
//
// Parity pins (tests/test_oracle_*.py):
//   random, uniform sphere  -- agreement fixture frame 0 (web/tests/fixtures/agreement.otj),
//                              bitwise against the reference library where it is built
//   plummer                 -- bitwise against the reference library; closed-form
//                              potential energy -3 pi M^2 / (32 a) within sampling noise
//   reference_acceleration  -- two-body closed form (0.75, -0.5), Plummer factor
//                              5*4/(4^2+3^2)^(3/2), reference library bitwise
//   direct (scalar, AVX2)   -- kepler-double.otj bitwise over 200 VV steps; agreement
//                              fixture energies; reference library bitwise
//   morton                  -- known answers of tests/solvers/morton_test.cpp:45-182;
//                              reference library bitwise (keys and permutation)
//   octree                  -- invariants of tests/solvers/octree_test.cpp; reference
//                              library bitwise (structure and moments)
//   barnes_hut              -- theta = 0 equals direct with pp = N(N-1); reference
//                              library bitwise (accelerations and counts)
//   integrators             -- kepler-double.otj bitwise; measured orders 2, 4 and 4 (RK4)
//   diagnostics             -- agreement.csv energies; Plummer and uniform-sphere
//                              closed forms
//   disc_collision          -- this project's own generator (no reference
//                              counterpart): mass, count, truncation and COM invariants
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <numbers>
#include <random>
#include <thread>
#include <vector>

#include "oracle_internal.hpp"

namespace oracle {

// ---------------------------------------------------------------------------
// Scalar direct kernel: direct_kernel_scalar.cpp:31-59.
template<typename R>
V3<R> accumulate_scalar(const Sources<R>& s, V3<R> t, std::size_t begin, std::size_t end,
                        R e2) noexcept {
    R ax = 0, ay = 0, az = 0;
    for (std::size_t j = begin; j < end; ++j) {
        const R dx = s.x[j] - t.x;
        const R dy = s.y[j] - t.y;
        const R dz = s.z[j] - t.z;
        const R r2 = (dx * dx) + (dy * dy) + (dz * dz);
        // softening.hpp:80-100: inverse = 1/sqrt(r2 + eps2), cubed as inv*inv*inv.
        const R inv = R{1} / std::sqrt(r2 + e2);
        const R f = s.m[j] * (inv * inv * inv);
        ax += f * dx;
        ay += f * dy;
        az += f * dz;
    }
    return {ax, ay, az};
}
template V3<double> accumulate_scalar<double>(const Sources<double>&, V3<double>, std::size_t,
                                              std::size_t, double) noexcept;
template V3<float> accumulate_scalar<float>(const Sources<float>&, V3<float>, std::size_t,
                                            std::size_t, float) noexcept;

namespace {

// ---------------------------------------------------------------------------
// Work-stealing parallel-for, restated from backend/work_stealing_executor.cpp:23-151
// and backend/partition.hpp: equal shares in cache-line grains, chunk =
// max(1, lines / (16 workers)) lines, claim from the front of the own share,
// steal from the back of the others'.
class Executor {
public:
    explicit Executor(unsigned workers) : workers_(std::max(1u, workers)) {}

    void run(std::size_t count, const std::function<void(std::size_t, std::size_t)>& task) {
        if (count == 0) return;
        constexpr std::size_t kGrain = 8;  // 64-byte line of FP64
        constexpr std::size_t kChunksPerWorker = 16;
        const unsigned w = workers_;
        if (w == 1) {
            task(0, count);
            return;
        }
        const std::size_t lines = (count + kGrain - 1) / kGrain;
        const std::size_t grain = std::max<std::size_t>(1, lines / (std::size_t{w} * kChunksPerWorker)) * kGrain;
        struct Range {
            std::mutex lock;
            std::size_t next = 0, end = 0;
        };
        std::vector<Range> ranges(w);
        for (unsigned k = 0; k < w; ++k) {
            const std::size_t share = lines / w, rem = lines % w;
            const std::size_t first = k * share + std::min<std::size_t>(k, rem);
            const std::size_t n_lines = share + (k < rem ? 1 : 0);
            ranges[k].next = std::min(count, first * kGrain);
            ranges[k].end = std::min(count, (first + n_lines) * kGrain);
        }
        auto drain = [&](unsigned self) {
            while (true) {
                std::size_t b = 0, e = 0;
                bool got = false;
                {
                    Range& r = ranges[self];
                    std::scoped_lock g{r.lock};
                    if (r.next < r.end) {
                        b = r.next;
                        e = std::min(r.next + grain, r.end);
                        r.next = e;
                        got = true;
                    }
                }
                for (unsigned off = 1; !got && off < w; ++off) {
                    Range& r = ranges[(self + off) % w];
                    std::scoped_lock g{r.lock};
                    if (r.next < r.end) {
                        const std::size_t take = std::min(grain, r.end - r.next);
                        e = r.end;
                        b = r.end - take;
                        r.end = b;
                        got = true;
                    }
                }
                if (!got) return;
                task(b, e);
            }
        };
        std::vector<std::thread> pool;
        pool.reserve(w - 1);
        for (unsigned k = 1; k < w; ++k) pool.emplace_back(drain, k);
        drain(0);
        for (auto& th : pool) th.join();
    }

private:
    unsigned workers_;
};

unsigned resolve_threads(unsigned threads) {
    if (threads != 0) return threads;
    const unsigned hw = std::thread::hardware_concurrency();
    return hw == 0 ? 1 : hw;
}

// ---------------------------------------------------------------------------
// Random source: core/random.cpp:12-47.
class Random {
public:
    explicit Random(std::uint64_t seed) : engine_(seed) {}
    double uniform() {
        const std::uint64_t bits = engine_() >> (64 - 53);
        return static_cast<double>(bits) * (1.0 / static_cast<double>(std::uint64_t{1} << 53));
    }
    double uniform(double lo, double hi) {
        const double f = uniform();
        return (lo * (1.0 - f)) + (hi * f);
    }
    V3<double> unit_vector() {
        const double h = uniform(-1, 1);
        const double az = uniform(0, 2 * std::numbers::pi_v<double>);
        const double r = std::sqrt(1.0 - (h * h));
        return {r * std::cos(az), r * std::sin(az), h};
    }
    // Box-Muller normal deviate for this project's own disc generator.
    double normal() {
        const double u1 = 1.0 - uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2 * std::numbers::pi_v<double> * u2);
    }

private:
    std::mt19937_64 engine_;
};

// Neumaier sum: core/diagnostics.cpp:36-59, reference_kernel.cpp.
class CompensatedSum {
public:
    void add(double v) noexcept {
        const double u = total_ + v;
        if (std::abs(total_) >= std::abs(v)) {
            correction_ += (total_ - u) + v;
        } else {
            correction_ += (v - u) + total_;
        }
        total_ = u;
    }
    double total() const noexcept { return total_ + correction_; }

private:
    double total_ = 0, correction_ = 0;
};

struct Particles {
    std::vector<double> x, y, z, vx, vy, vz, m;
    void add(V3<double> p, V3<double> v, double mass) {
        x.push_back(p.x);
        y.push_back(p.y);
        z.push_back(p.z);
        vx.push_back(v.x);
        vy.push_back(v.y);
        vz.push_back(v.z);
        m.push_back(mass);
    }
    std::size_t size() const { return m.size(); }
};

// centre_of_mass_frame.cpp with diagnostics.cpp centre_of_mass / linear_momentum.
void move_to_centre_of_mass_frame(Particles& p) {
    CompensatedSum mass;
    for (double mi : p.m) mass.add(mi);
    const double total = mass.total();
    CompensatedSum wx, wy, wz, px, py, pz;
    for (std::size_t i = 0; i < p.size(); ++i) {
        wx.add(p.x[i] * p.m[i]);
        wy.add(p.y[i] * p.m[i]);
        wz.add(p.z[i] * p.m[i]);
    }
    for (std::size_t i = 0; i < p.size(); ++i) {
        px.add(p.vx[i] * p.m[i]);
        py.add(p.vy[i] * p.m[i]);
        pz.add(p.vz[i] * p.m[i]);
    }
    V3<double> c{}, d{};
    if (total != 0) {
        c = {wx.total() / total, wy.total() / total, wz.total() / total};
        d = {px.total() / total, py.total() / total, pz.total() / total};
    }
    for (std::size_t i = 0; i < p.size(); ++i) {
        p.x[i] = p.x[i] - c.x;
        p.y[i] = p.y[i] - c.y;
        p.z[i] = p.z[i] - c.z;
        p.vx[i] = p.vx[i] - d.x;
        p.vy[i] = p.vy[i] - d.y;
        p.vz[i] = p.vz[i] - d.z;
    }
}

void copy_out(const Particles& p, double* x, double* y, double* z, double* vx, double* vy,
              double* vz, double* m) {
    const std::size_t n = p.size();
    std::memcpy(x, p.x.data(), n * 8);
    std::memcpy(y, p.y.data(), n * 8);
    std::memcpy(z, p.z.data(), n * 8);
    std::memcpy(vx, p.vx.data(), n * 8);
    std::memcpy(vy, p.vy.data(), n * 8);
    std::memcpy(vz, p.vz.data(), n * 8);
    std::memcpy(m, p.m.data(), n * 8);
}

V3<double> scale(double s, V3<double> v) { return {v.x * s, v.y * s, v.z * s}; }

// ---------------------------------------------------------------------------
// Morton: solvers/morton.cpp.
constexpr unsigned kBits = 21;
constexpr std::uint32_t kGrid = std::uint32_t{1} << kBits;

struct Cube {
    V3<double> origin;
    double size = 0;
    V3<double> centre() const {  // morton.hpp:93-95
        return {origin.x + size / 2, origin.y + size / 2, origin.z + size / 2};
    }
};

Cube bounding_cube(std::size_t n, const double* x, const double* y, const double* z) {
    if (n == 0) return Cube{};
    V3<double> lo{x[0], y[0], z[0]}, hi = lo;
    for (std::size_t i = 1; i < n; ++i) {
        lo.x = std::min(lo.x, x[i]);
        lo.y = std::min(lo.y, y[i]);
        lo.z = std::min(lo.z, z[i]);
        hi.x = std::max(hi.x, x[i]);
        hi.y = std::max(hi.y, y[i]);
        hi.z = std::max(hi.z, z[i]);
    }
    double size = std::max({hi.x - lo.x, hi.y - lo.y, hi.z - lo.z});
    if (!(size > 0)) size = 1;
    const V3<double> c{(lo.x + hi.x) / 2, (lo.y + hi.y) / 2, (lo.z + hi.z) / 2};
    return Cube{{c.x - size / 2, c.y - size / 2, c.z - size / 2}, size};
}

std::uint64_t spread(std::uint32_t v) {
    std::uint64_t s = v & 0x1fffffu;
    s = (s | (s << 32)) & 0x1f00000000ffffull;
    s = (s | (s << 16)) & 0x1f0000ff0000ffull;
    s = (s | (s << 8)) & 0x100f00f00f00f00full;
    s = (s | (s << 4)) & 0x10c30c30c30c30c3ull;
    s = (s | (s << 2)) & 0x1249249249249249ull;
    return s;
}

std::uint32_t to_grid(double coordinate, double origin, double scale) {
    const double cell = (coordinate - origin) * scale;
    if (!(cell > 0)) return 0;
    if (cell >= static_cast<double>(kGrid)) return kGrid - 1;
    return static_cast<std::uint32_t>(cell);
}

std::uint64_t morton(double x, double y, double z, const Cube& c) {
    if (!(c.size > 0)) return 0;
    const double s = static_cast<double>(kGrid) / c.size;
    return (spread(to_grid(x, c.origin.x, s)) << 2) | (spread(to_grid(y, c.origin.y, s)) << 1) |
           spread(to_grid(z, c.origin.z, s));
}

unsigned octant(std::uint64_t code, unsigned level) {
    if (level == 0 || level > kBits) return 0;
    return static_cast<unsigned>((code >> (3 * (kBits - level))) & 7u);
}

struct Key {
    std::uint64_t code;
    std::uint64_t index;
    bool operator<(const Key& o) const {
        return code != o.code ? code < o.code : index < o.index;
    }
};

std::vector<Key> morton_order(std::size_t n, const double* x, const double* y, const double* z,
                              Cube& cube) {
    cube = bounding_cube(n, x, y, z);
    std::vector<Key> keys(n);
    for (std::size_t i = 0; i < n; ++i) keys[i] = Key{morton(x[i], y[i], z[i], cube), i};
    std::sort(keys.begin(), keys.end());
    return keys;
}

}  // namespace

// ---------------------------------------------------------------------------
// Octree: solvers/octree.cpp (serial emit, which the reference's parallel
// build reproduces exactly: octree_test.cpp:453-492).
struct Quad {
    double xx = 0, xy = 0, xz = 0, yy = 0, yz = 0, zz = 0;
};
struct Node {
    double mass = 0;
    V3<double> com;
    double acc_r2 = 0;
    std::size_t next = 0, first = 0, count = 0;
};

}  // namespace oracle

struct orc_tree {
    std::vector<oracle::Node> nodes;
    std::vector<oracle::Quad> quads;
    std::size_t leaves = 0;
    unsigned depth = 0;
};

namespace oracle {
namespace {

struct TreeBuilder {
    const double *x, *y, *z, *m;
    const std::uint64_t* codes;
    double theta;
    std::size_t cap;
    bool quadrupole;
    orc_tree* out;

    static V3<double> child_centre(V3<double> c, double size, unsigned oct) {  // octree.cpp:41-47
        const double q = size / 4;
        return {c.x + ((oct & 4u) != 0 ? q : -q), c.y + ((oct & 2u) != 0 ? q : -q),
                c.z + ((oct & 1u) != 0 ? q : -q)};
    }
    std::size_t octant_end(std::size_t b, std::size_t e, unsigned level, unsigned oct) const {
        // octree.cpp:56-66: partition point of octant(code, level) <= oct.
        std::size_t lo = b, hi = e;
        while (lo < hi) {
            const std::size_t mid = lo + (hi - lo) / 2;
            if (octant(codes[mid], level) <= oct) lo = mid + 1; else hi = mid;
        }
        return lo;
    }
    static void displaced(Quad& q, double mass, V3<double> d) {  // octree.cpp:75-84
        const double trace = mass * ((d.x * d.x) + (d.y * d.y) + (d.z * d.z));
        q.xx += (3 * mass * d.x * d.x) - trace;
        q.xy += 3 * mass * d.x * d.y;
        q.xz += 3 * mass * d.x * d.z;
        q.yy += (3 * mass * d.y * d.y) - trace;
        q.yz += 3 * mass * d.y * d.z;
        q.zz += (3 * mass * d.z * d.z) - trace;
    }
    void leaf_moments(std::size_t self) {  // octree.cpp:91-135
        Node& nd = out->nodes[self];
        const std::size_t b = nd.first, e = b + nd.count;
        double mass = 0;
        V3<double> w{};
        for (std::size_t i = b; i < e; ++i) {
            mass += m[i];
            w.x += x[i] * m[i];
            w.y += y[i] * m[i];
            w.z += z[i] * m[i];
        }
        nd.mass = mass;
        nd.com = mass > 0 ? V3<double>{w.x / mass, w.y / mass, w.z / mass} : V3<double>{x[b], y[b], z[b]};
        if (!quadrupole) return;
        Quad q{};
        for (std::size_t i = b; i < e; ++i) {
            displaced(q, m[i], {x[i] - nd.com.x, y[i] - nd.com.y, z[i] - nd.com.z});
        }
        out->quads[self] = q;
    }
    void gather(std::size_t self) {  // octree.cpp:142-185
        auto& nodes = out->nodes;
        const std::size_t end = nodes[self].next;
        double mass = 0;
        V3<double> w{}, mean{};
        std::size_t children = 0;
        for (std::size_t c = self + 1; c < end; c = nodes[c].next) {
            mass += nodes[c].mass;
            w.x += nodes[c].com.x * nodes[c].mass;
            w.y += nodes[c].com.y * nodes[c].mass;
            w.z += nodes[c].com.z * nodes[c].mass;
            mean.x += nodes[c].com.x;
            mean.y += nodes[c].com.y;
            mean.z += nodes[c].com.z;
            ++children;
        }
        nodes[self].mass = mass;
        if (mass > 0) {
            nodes[self].com = {w.x / mass, w.y / mass, w.z / mass};
        } else {
            const double k = static_cast<double>(children == 0 ? 1 : children);
            nodes[self].com = {mean.x / k, mean.y / k, mean.z / k};
        }
        if (!quadrupole) return;
        Quad q{};
        for (std::size_t c = self + 1; c < end; c = nodes[c].next) {
            const Quad& cq = out->quads[c];
            q.xx += cq.xx;
            q.xy += cq.xy;
            q.xz += cq.xz;
            q.yy += cq.yy;
            q.yz += cq.yz;
            q.zz += cq.zz;
            displaced(q, nodes[c].mass,
                      {nodes[c].com.x - nodes[self].com.x, nodes[c].com.y - nodes[self].com.y,
                       nodes[c].com.z - nodes[self].com.z});
        }
        out->quads[self] = q;
    }
    double acceptance(double size, double offset) const {  // octree.cpp:374-387
        if (!(theta > 0)) return std::numeric_limits<double>::infinity();
        const double r = (size / theta) + offset;
        return r * r;
    }
    void emit(std::size_t b, std::size_t e, unsigned level, V3<double> centre, double size) {
        // octree.cpp:320-372
        const std::size_t self = out->nodes.size();
        out->nodes.emplace_back();
        if (quadrupole) out->quads.emplace_back();
        if ((e - b <= cap) || level >= kBits) {
            out->nodes[self].first = b;
            out->nodes[self].count = e - b;
            leaf_moments(self);
            ++out->leaves;
            out->depth = std::max(out->depth, level);
            out->nodes[self].next = out->nodes.size();
        } else {
            std::size_t cb = b;
            for (unsigned oct = 0; oct < 8; ++oct) {
                const std::size_t ce = octant_end(cb, e, level + 1, oct);
                if (ce > cb) emit(cb, ce, level + 1, child_centre(centre, size, oct), size / 2);
                cb = ce;
            }
            out->nodes[self].next = out->nodes.size();
            gather(self);
        }
        const V3<double> c = out->nodes[self].com;
        const V3<double> d{c.x - centre.x, c.y - centre.y, c.z - centre.z};
        const double offset = std::sqrt((d.x * d.x) + (d.y * d.y) + (d.z * d.z));
        out->nodes[self].acc_r2 = acceptance(size, offset);
    }
};

// tree_walk.hpp:129-139 monopole and :156-181 quadrupole terms, no G.
V3<double> monopole(V3<double> off, double mass, double e2) {
    const double r2 = (off.x * off.x) + (off.y * off.y) + (off.z * off.z);
    const double inv = 1.0 / std::sqrt(r2 + e2);
    const double f = mass * (inv * inv * inv);
    return {off.x * f, off.y * f, off.z * f};
}
V3<double> quadrupole_term(V3<double> d, const Quad& q, double e2) {
    const double r2 = (d.x * d.x) + (d.y * d.y) + (d.z * d.z);
    const double inv = 1.0 / std::sqrt(r2 + e2);
    const double i2 = inv * inv;
    const double i5 = i2 * i2 * inv;
    const double i7 = i5 * i2;
    const V3<double> c{(q.xx * d.x) + (q.xy * d.y) + (q.xz * d.z),
                       (q.xy * d.x) + (q.yy * d.y) + (q.yz * d.z),
                       (q.xz * d.x) + (q.yz * d.y) + (q.zz * d.z)};
    const double quadratic = (d.x * c.x) + (d.y * c.y) + (d.z * c.z);
    const double s = 2.5 * quadratic * i7;
    const V3<double> a{c.x * -i5, c.y * -i5, c.z * -i5};
    return {a.x + (d.x * s), a.y + (d.y * s), a.z + (d.z * s)};
}

using Accumulate = V3<double> (*)(const Sources<double>&, V3<double>, std::size_t, std::size_t,
                                  double) noexcept;

V3<double> scalar_kernel(const Sources<double>& s, V3<double> t, std::size_t b, std::size_t e,
                         double e2) noexcept {
    return accumulate_scalar<double>(s, t, b, e, e2);
}
V3<double> avx2_kernel(const Sources<double>& s, V3<double> t, std::size_t b, std::size_t e,
                       double e2) noexcept {
    return accumulate_avx2(s, t, b, e, e2);
}

V3<double> add(V3<double> a, V3<double> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }

// tree_walk.cpp:20-83
V3<double> walk(const orc_tree& tree, const Sources<double>& s, std::size_t target, double e2,
                Accumulate acc, std::uint64_t& pp, std::uint64_t& pc) {
    const auto& nodes = tree.nodes;
    const bool quad = !tree.quads.empty();
    const V3<double> p{s.x[target], s.y[target], s.z[target]};
    V3<double> a{};
    std::size_t node = 0;
    while (node < nodes.size()) {
        const Node& c = nodes[node];
        const V3<double> off{c.com.x - p.x, c.com.y - p.y, c.com.z - p.z};
        const double r2 = (off.x * off.x) + (off.y * off.y) + (off.z * off.z);
        if (r2 >= c.acc_r2) {
            a = add(a, monopole(off, c.mass, e2));
            if (quad) a = add(a, quadrupole_term(off, tree.quads[node], e2));
            ++pc;
            node = c.next;
        } else if (c.count > 0) {
            const std::size_t b = c.first, e = b + c.count;
            if (target >= b && target < e) {
                a = add(a, acc(s, p, b, target, e2));
                a = add(a, acc(s, p, target + 1, e, e2));
                pp += c.count - 1;
            } else {
                a = add(a, acc(s, p, b, e, e2));
                pp += c.count;
            }
            node = c.next;
        } else {
            ++node;
        }
    }
    return a;
}

Accumulate pick_kernel(int kernel) { return kernel == 1 ? avx2_kernel : scalar_kernel; }

// Direct solver over one precision: direct_solver.cpp:20-80.
template<typename R>
void direct_impl(std::size_t n, const R* x, const R* y, const R* z, const R* m, R e2, R G,
                 int kernel, unsigned threads, std::size_t tb, std::size_t te, R* ax, R* ay, R* az) {
    const Sources<R> s{x, y, z, m};
    Executor ex(resolve_threads(threads));
    ex.run(te - tb, [&](std::size_t b, std::size_t e) {
        for (std::size_t k = b; k < e; ++k) {
            const std::size_t i = tb + k;
            const V3<R> t{x[i], y[i], z[i]};
            V3<R> before, after;
            if (kernel == 1) {
                before = accumulate_avx2(s, t, 0, i, e2);
                after = accumulate_avx2(s, t, i + 1, n, e2);
            } else {
                before = accumulate_scalar<R>(s, t, 0, i, e2);
                after = accumulate_scalar<R>(s, t, i + 1, n, e2);
            }
            // G * (before + after): Vec3 sum then scalar times vector (vec3.hpp:78-101).
            ax[i] = G * (before.x + after.x);
            ay[i] = G * (before.y + after.y);
            az[i] = G * (before.z + after.z);
        }
    });
}

// barnes_hut_solver.cpp:41-130
void barnes_hut(std::size_t n, const double* x, const double* y, const double* z,
                const double* m, double e2, double G, double theta, std::size_t cap, bool quad,
                int kernel, unsigned threads, double* ax, double* ay, double* az,
                std::uint64_t counts[2]) {
    Cube cube;
    const std::vector<Key> keys = morton_order(n, x, y, z, cube);
    std::vector<double> sx(n), sy(n), sz(n), sm(n);
    std::vector<std::uint64_t> codes(n);
    for (std::size_t i = 0; i < n; ++i) {
        const std::size_t src = keys[i].index;
        sx[i] = x[src];
        sy[i] = y[src];
        sz[i] = z[src];
        sm[i] = m[src];
        codes[i] = keys[i].code;
    }
    const double origin[3] = {cube.origin.x, cube.origin.y, cube.origin.z};
    orc_tree* tree = orc_tree_build(n, sx.data(), sy.data(), sz.data(), sm.data(), codes.data(),
                                    origin, cube.size, theta, cap, quad ? 1 : 0);
    const Sources<double> s{sx.data(), sy.data(), sz.data(), sm.data()};
    const Accumulate acc = pick_kernel(kernel);
    std::atomic<std::uint64_t> pp{0}, pc{0};
    Executor ex(resolve_threads(threads));
    ex.run(n, [&](std::size_t b, std::size_t e) {
        std::uint64_t lpp = 0, lpc = 0;
        for (std::size_t i = b; i < e; ++i) {
            const V3<double> a = walk(*tree, s, i, e2, acc, lpp, lpc);
            const std::size_t dst = keys[i].index;
            ax[dst] = G * a.x;
            ay[dst] = G * a.y;
            az[dst] = G * a.z;
        }
        pp.fetch_add(lpp, std::memory_order_relaxed);
        pc.fetch_add(lpc, std::memory_order_relaxed);
    });
    orc_tree_free(tree);
    counts[0] = pp.load();
    counts[1] = pc.load();
}

void field_eval(std::size_t n, const double* x, const double* y, const double* z, const double* m,
                const orc_field& f, double* ax, double* ay, double* az) {
    if (f.solver == 1) {
        std::uint64_t counts[2];
        barnes_hut(n, x, y, z, m, f.softening_squared, f.G, f.opening_angle,
                   std::max<std::size_t>(1, f.leaf_capacity), false, 1, f.threads, ax, ay, az, counts);
    } else {
        direct_impl<double>(n, x, y, z, m, f.softening_squared, f.G, f.solver == 2 ? 1 : 0,
                            f.threads, 0, n, ax, ay, az);
    }
}

// velocity_verlet.cpp:20-72
void vv_step(std::size_t n, double* x, double* y, double* z, double* vx, double* vy, double* vz,
             double* ax, double* ay, double* az, const double* m, const orc_field& f, double dt) {
    const double half = dt / 2;
    for (std::size_t i = 0; i < n; ++i) vx[i] += half * ax[i];
    for (std::size_t i = 0; i < n; ++i) vy[i] += half * ay[i];
    for (std::size_t i = 0; i < n; ++i) vz[i] += half * az[i];
    for (std::size_t i = 0; i < n; ++i) x[i] += dt * vx[i];
    for (std::size_t i = 0; i < n; ++i) y[i] += dt * vy[i];
    for (std::size_t i = 0; i < n; ++i) z[i] += dt * vz[i];
    field_eval(n, x, y, z, m, f, ax, ay, az);
    for (std::size_t i = 0; i < n; ++i) vx[i] += half * ax[i];
    for (std::size_t i = 0; i < n; ++i) vy[i] += half * ay[i];
    for (std::size_t i = 0; i < n; ++i) vz[i] += half * az[i];
}

// runge_kutta4.cpp:64-139: classical RK4 on (x, v) with stage sums; four
// evaluations per step, the last restoring a(x) at the final positions.
void rk4_step(std::size_t n, double* x, double* y, double* z, double* vx, double* vy, double* vz,
              double* ax, double* ay, double* az, const double* m, const orc_field& f, double dt) {
    double* pos[3] = {x, y, z};
    double* vel[3] = {vx, vy, vz};
    double* acc[3] = {ax, ay, az};
    std::vector<double> p0[3], v0[3], sp[3], sv[3];
    for (int c = 0; c < 3; ++c) {
        p0[c].assign(pos[c], pos[c] + n);
        v0[c].assign(vel[c], vel[c] + n);
        sp[c].assign(vel[c], vel[c] + n);
        sv[c].assign(acc[c], acc[c] + n);
    }
    auto stage = [&](double scale, double weight) {
        for (int c = 0; c < 3; ++c) {
            for (std::size_t i = 0; i < n; ++i) {
                const double stage_velocity = vel[c][i];
                pos[c][i] = p0[c][i] + (scale * stage_velocity);
                vel[c][i] = v0[c][i] + (scale * acc[c][i]);
            }
        }
        field_eval(n, x, y, z, m, f, ax, ay, az);
        for (int c = 0; c < 3; ++c) {
            for (std::size_t i = 0; i < n; ++i) sp[c][i] += weight * vel[c][i];
        }
        for (int c = 0; c < 3; ++c) {
            for (std::size_t i = 0; i < n; ++i) sv[c][i] += weight * acc[c][i];
        }
    };
    const double half = dt / 2;
    stage(half, 2);
    stage(half, 2);
    stage(dt, 1);
    const double sixth = dt / 6;
    for (int c = 0; c < 3; ++c) {
        for (std::size_t i = 0; i < n; ++i) pos[c][i] = p0[c][i] + (sixth * sp[c][i]);
    }
    for (int c = 0; c < 3; ++c) {
        for (std::size_t i = 0; i < n; ++i) vel[c][i] = v0[c][i] + (sixth * sv[c][i]);
    }
    field_eval(n, x, y, z, m, f, ax, ay, az);
}

// yoshida4.cpp:24-41
constexpr double kCubeRootOfTwo = 1.2599210498948731647672106;
constexpr double kOuter = 1.0 / (2 - kCubeRootOfTwo);
constexpr double kInner = -kCubeRootOfTwo / (2 - kCubeRootOfTwo);

// ---------------------------------------------------------------------------
// Two-disc collision generator (this project's own model).
struct DiscModel {
    double mass, rd, zd, rmax, bulge_mass, bulge_a;
    // Enclosed-mass circular speed: spherical Hernquist bulge plus the disc
    // mass inside cylindrical radius R, treated as if spherical.
    double disc_mass_fraction(double R) const {
        const double norm = 1 - (1 + rmax / rd) * std::exp(-rmax / rd);
        const double u = R / rd;
        return std::min(1.0, (1 - (1 + u) * std::exp(-u)) / norm);
    }
    double bulge_mass_fraction(double r) const {
        const double f = (r * r) / ((r + bulge_a) * (r + bulge_a));
        return std::min(1.0, f);
    }
    double circular_speed(double r, double G) const {
        if (!(r > 0)) return 0;
        const double enclosed = (mass - bulge_mass) * disc_mass_fraction(r) +
                                bulge_mass * bulge_mass_fraction(r);
        return std::sqrt(G * enclosed / r);
    }
    double sample_disc_radius(double u) const {  // bisection on the truncated CDF
        double lo = 0, hi = rmax;
        for (int it = 0; it < 80; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (disc_mass_fraction(mid) < u) lo = mid; else hi = mid;
        }
        return 0.5 * (lo + hi);
    }
};

void make_galaxy(Particles& out, std::size_t count, const DiscModel& g, double bulge_fraction,
                 double dispersion, double inclined_fraction, double inclination, Random& rnd,
                 double G) {
    const double pm = g.mass / static_cast<double>(count);
    const std::size_t n_bulge = static_cast<std::size_t>(std::llround(bulge_fraction * static_cast<double>(count)));
    for (std::size_t k = 0; k < count; ++k) {
        if (k < n_bulge) {
            // Hernquist radius by inverse CDF M(<r) = M r^2/(r+a)^2, truncated at 0.99 of the mass.
            const double u = rnd.uniform(0, 0.99);
            const double su = std::sqrt(u);
            const double r = g.bulge_a * su / (1 - su);
            const V3<double> dir = rnd.unit_vector();
            const double sigma = g.circular_speed(r, G) / std::sqrt(3.0);
            const double v1 = rnd.normal();
            const double v2 = rnd.normal();
            const double v3 = rnd.normal();
            out.add(scale(r, dir), {sigma * v1, sigma * v2, sigma * v3}, pm);
        } else {
            const double u = rnd.uniform();
            const double R = g.sample_disc_radius(u);
            const double phi = rnd.uniform(0, 2 * std::numbers::pi_v<double>);
            const double w = rnd.uniform(-1, 1);
            const double zz = g.zd * std::atanh(std::clamp(w, -1 + 1e-15, 1 - 1e-15));  // sech^2 profile
            const double vc = g.circular_speed(R, G);
            const double n1 = rnd.normal();
            const double n2 = rnd.normal();
            const double n3 = rnd.normal();
            const double c = std::cos(phi), s = std::sin(phi);
            V3<double> pos{R * c, R * s, zz};
            V3<double> vel{-vc * s + dispersion * vc * n1, vc * c + dispersion * vc * n2,
                           dispersion * vc * n3};
            const double tilt = rnd.uniform();
            if (tilt < inclined_fraction) {  // rotate about the x axis
                const double ci = std::cos(inclination), si = std::sin(inclination);
                pos = {pos.x, ci * pos.y - si * pos.z, si * pos.y + ci * pos.z};
                vel = {vel.x, ci * vel.y - si * vel.z, si * vel.y + ci * vel.z};
            }
            out.add(pos, vel, pm);
        }
    }
}

}  // namespace
}  // namespace oracle

using namespace oracle;

extern "C" {

int orc_random_uniform(uint64_t seed, uint64_t count, double* out) {
    Random r(seed);
    for (uint64_t i = 0; i < count; ++i) out[i] = r.uniform();
    return 0;
}

int orc_plummer(uint64_t count, double M, double a, double cutoff, uint64_t seed, double* x,
                double* y, double* z, double* vx, double* vy, double* vz, double* m) {
    return orc_plummer_ordered(count, M, a, cutoff, seed, 0, x, y, z, vx, vy, vz, m);
}

int orc_plummer_ordered(uint64_t count, double M, double a, double cutoff, uint64_t seed,
                        int velocity_direction_first, double* x, double* y, double* z, double* vx,
                        double* vy, double* vz, double* m) {
    if (!(M > 0) || !(a > 0) || !(cutoff > 0) || !(cutoff < 1)) return 1;
    if (count == 0) return 0;
    Random rnd(seed);
    Particles p;
    const double pm = M / static_cast<double>(count);
    for (uint64_t k = 0; k < count; ++k) {
        const double mf = rnd.uniform(0, cutoff);
        const double r = a / std::sqrt(std::pow(mf, -2.0 / 3) - 1);  // plummer.cpp:43-45
        double q;
        while (true) {  // plummer.cpp:52-87, envelope 0.1
            q = rnd.uniform();
            const double c = 1 - (q * q);
            const double density = q * q * c * c * c * std::sqrt(c);
            if (0.1 * rnd.uniform() < density) break;
        }
        const double vesc = std::sqrt(2 * 1.0 * M / std::sqrt((r * r) + (a * a)));
        const double speed = q * vesc;
        // plummer.cpp:122 draws both directions inside one argument list:
        // left to right (Clang) the position's first, right to left (GCC) the
        // velocity's.
        V3<double> pd, vd;
        if (velocity_direction_first != 0) {
            vd = rnd.unit_vector();
            pd = rnd.unit_vector();
        } else {
            pd = rnd.unit_vector();
            vd = rnd.unit_vector();
        }
        p.add(scale(r, pd), scale(speed, vd), pm);
    }
    move_to_centre_of_mass_frame(p);
    copy_out(p, x, y, z, vx, vy, vz, m);
    return 0;
}

int orc_uniform_sphere(uint64_t count, double M, double R, uint64_t seed, double* x, double* y,
                       double* z, double* vx, double* vy, double* vz, double* m) {
    if (!(M > 0) || !(R > 0)) return 1;
    if (count == 0) return 0;
    Random rnd(seed);
    Particles p;
    const double pm = M / static_cast<double>(count);
    for (uint64_t k = 0; k < count; ++k) {
        const double r = R * std::cbrt(rnd.uniform());
        p.add(scale(r, rnd.unit_vector()), V3<double>{}, pm);
    }
    move_to_centre_of_mass_frame(p);
    copy_out(p, x, y, z, vx, vy, vz, m);
    return 0;
}

int orc_disc_collision(const orc_disc_collision_params* p, double* x, double* y, double* z,
                       double* vx, double* vy, double* vz, double* m) {
    if (p == nullptr || p->count < 2 || !(p->mass_ratio > 0) || !(p->total_mass > 0)) return 1;
    const double G = 1.0;
    Random rnd(p->seed);
    const double M1 = p->total_mass * p->mass_ratio / (1 + p->mass_ratio);
    const double M2 = p->total_mass - M1;
    const std::size_t n1 = static_cast<std::size_t>(std::llround(static_cast<double>(p->count) * M1 / p->total_mass));
    const std::size_t n2 = p->count - n1;
    Particles all;
    const DiscModel g1{M1, p->scale_length[0], p->scale_height, p->truncation * p->scale_length[0],
                       p->bulge_fraction * M1, p->bulge_scale};
    const DiscModel g2{M2, p->scale_length[1], p->scale_height, p->truncation * p->scale_length[1],
                       p->bulge_fraction * M2, p->bulge_scale};
    Particles a, b;
    make_galaxy(a, n1, g1, p->bulge_fraction, p->dispersion, p->inclined_fraction, p->inclination, rnd, G);
    make_galaxy(b, n2, g2, p->bulge_fraction, p->dispersion, p->inclined_fraction, p->inclination, rnd, G);
    // Parabolic relative orbit: energy 0, h = sqrt(2 G M q).
    const double Mt = M1 + M2, d = p->separation, q = p->pericentre;
    const double v = std::sqrt(2 * G * Mt / d);
    const double h = std::sqrt(2 * G * Mt * q);
    const double vt = h / d;
    const double vr = -std::sqrt(std::max(0.0, v * v - vt * vt));
    const V3<double> rrel{d, 0, 0}, vrel{vr, vt, 0};
    const double f1 = M2 / Mt, f2 = M1 / Mt;
    for (std::size_t i = 0; i < a.size(); ++i) {
        all.add({a.x[i] - f1 * rrel.x, a.y[i] - f1 * rrel.y, a.z[i] - f1 * rrel.z},
                {a.vx[i] - f1 * vrel.x, a.vy[i] - f1 * vrel.y, a.vz[i] - f1 * vrel.z}, a.m[i]);
    }
    for (std::size_t i = 0; i < b.size(); ++i) {
        all.add({b.x[i] + f2 * rrel.x, b.y[i] + f2 * rrel.y, b.z[i] + f2 * rrel.z},
                {b.vx[i] + f2 * vrel.x, b.vy[i] + f2 * vrel.y, b.vz[i] + f2 * vrel.z}, b.m[i]);
    }
    move_to_centre_of_mass_frame(all);
    copy_out(all, x, y, z, vx, vy, vz, m);
    return 0;
}

int orc_reference_acceleration(uint64_t n, const double* x, const double* y, const double* z,
                               const double* m, double e2, double G, const uint64_t* targets,
                               uint64_t tcount, double* ax, double* ay, double* az,
                               unsigned threads) {
    const uint64_t total = targets == nullptr ? n : tcount;
    Executor ex(resolve_threads(threads));
    ex.run(total, [&](std::size_t b, std::size_t e) {
        for (std::size_t k = b; k < e; ++k) {
            const std::size_t t = targets == nullptr ? k : targets[k];
            if (t >= n) continue;
            const double tx = x[t], ty = y[t], tz = z[t];
            CompensatedSum sx, sy, sz;
            for (std::size_t j = 0; j < n; ++j) {
                if (j == t) continue;
                const double dx = x[j] - tx, dy = y[j] - ty, dz = z[j] - tz;
                const double r2 = (dx * dx) + (dy * dy) + (dz * dz) + e2;
                const double inv = 1.0 / std::sqrt(r2);
                const double f = m[j] * inv * inv * inv;
                sx.add(f * dx);
                sy.add(f * dy);
                sz.add(f * dz);
            }
            ax[k] = G * sx.total();
            ay[k] = G * sy.total();
            az[k] = G * sz.total();
        }
    });
    return 0;
}

int orc_direct(uint64_t n, const double* x, const double* y, const double* z, const double* m,
               double e2, double G, int kernel, unsigned threads, uint64_t tb, uint64_t te,
               double* ax, double* ay, double* az) {
    if (tb > te || te > n) return 1;
    direct_impl<double>(n, x, y, z, m, e2, G, kernel, threads, tb, te, ax, ay, az);
    return 0;
}

int orc_direct_f32(uint64_t n, const float* x, const float* y, const float* z, const float* m,
                   float e2, float G, int kernel, unsigned threads, uint64_t tb, uint64_t te,
                   float* ax, float* ay, float* az) {
    if (tb > te || te > n) return 1;
    direct_impl<float>(n, x, y, z, m, e2, G, kernel, threads, tb, te, ax, ay, az);
    return 0;
}

int orc_bounding_cube(uint64_t n, const double* x, const double* y, const double* z,
                      double origin[3], double* size) {
    const Cube c = bounding_cube(n, x, y, z);
    origin[0] = c.origin.x;
    origin[1] = c.origin.y;
    origin[2] = c.origin.z;
    *size = c.size;
    return 0;
}

uint64_t orc_spread_bits(uint32_t value) { return spread(value); }

uint64_t orc_morton_interleave(uint32_t x, uint32_t y, uint32_t z) {
    return (spread(x) << 2) | (spread(y) << 1) | spread(z);
}

unsigned orc_morton_octant(uint64_t code, unsigned level) { return octant(code, level); }

int orc_morton_codes(uint64_t n, const double* x, const double* y, const double* z,
                     const double origin[3], double size, uint64_t* codes) {
    Cube c{{origin[0], origin[1], origin[2]}, size};
    for (uint64_t i = 0; i < n; ++i) codes[i] = morton(x[i], y[i], z[i], c);
    return 0;
}

int orc_morton_order(uint64_t n, const double* x, const double* y, const double* z,
                     double origin[3], double* size, uint64_t* codes, uint64_t* index) {
    Cube c;
    const std::vector<Key> keys = morton_order(n, x, y, z, c);
    for (uint64_t i = 0; i < n; ++i) {
        codes[i] = keys[i].code;
        index[i] = keys[i].index;
    }
    origin[0] = c.origin.x;
    origin[1] = c.origin.y;
    origin[2] = c.origin.z;
    *size = c.size;
    return 0;
}

orc_tree* orc_tree_build(uint64_t n, const double* sx, const double* sy, const double* sz,
                         const double* sm, const uint64_t* codes, const double origin[3],
                         double size, double theta, uint64_t cap, int quadrupole) {
    auto* tree = new orc_tree();
    if (n == 0) return tree;
    // octree.cpp:189-194 parameter correction.
    theta = std::clamp(theta, 0.0, 1.0);
    cap = std::max<uint64_t>(cap, 1);
    TreeBuilder b{sx, sy, sz, sm, codes, theta, cap, quadrupole != 0, tree};
    const Cube c{{origin[0], origin[1], origin[2]}, size};
    b.emit(0, n, 0, c.centre(), size);
    return tree;
}

uint64_t orc_tree_node_count(const orc_tree* t) { return t->nodes.size(); }
uint64_t orc_tree_leaf_count(const orc_tree* t) { return t->leaves; }
unsigned orc_tree_depth(const orc_tree* t) { return t->depth; }

void orc_tree_nodes(const orc_tree* t, double* mass, double* cx, double* cy, double* cz,
                    double* acc, uint64_t* next, uint64_t* first, uint64_t* count, double* quad) {
    for (std::size_t i = 0; i < t->nodes.size(); ++i) {
        const Node& nd = t->nodes[i];
        if (mass) mass[i] = nd.mass;
        if (cx) cx[i] = nd.com.x;
        if (cy) cy[i] = nd.com.y;
        if (cz) cz[i] = nd.com.z;
        if (acc) acc[i] = nd.acc_r2;
        if (next) next[i] = nd.next;
        if (first) first[i] = nd.first;
        if (count) count[i] = nd.count;
        if (quad && !t->quads.empty()) {
            const Quad& q = t->quads[i];
            double* o = quad + 6 * i;
            o[0] = q.xx; o[1] = q.xy; o[2] = q.xz; o[3] = q.yy; o[4] = q.yz; o[5] = q.zz;
        }
    }
}

void orc_tree_free(orc_tree* t) { delete t; }

int orc_barnes_hut(uint64_t n, const double* x, const double* y, const double* z, const double* m,
                   double e2, double G, double theta, uint64_t cap, int quadrupole, int kernel,
                   unsigned threads, double* ax, double* ay, double* az, uint64_t counts[2]) {
    barnes_hut(n, x, y, z, m, e2, G, theta, std::max<uint64_t>(cap, 1), quadrupole != 0, kernel,
               threads, ax, ay, az, counts);
    return 0;
}

int orc_barnes_hut_fp32_counts(uint64_t n, const double* x, const double* y, const double* z,
                               const double* m, double theta, uint64_t cap, unsigned threads,
                               uint64_t counts[2]) {
    counts[0] = counts[1] = 0;
    if (n == 0) return 0;
    Cube cube;
    const std::vector<Key> keys = morton_order(n, x, y, z, cube);
    std::vector<double> sx(n), sy(n), sz(n), sm(n);
    std::vector<std::uint64_t> codes(n);
    for (std::size_t i = 0; i < n; ++i) {
        const std::size_t src = keys[i].index;
        sx[i] = x[src];
        sy[i] = y[src];
        sz[i] = z[src];
        sm[i] = m[src];
        codes[i] = keys[i].code;
    }
    const double origin[3] = {cube.origin.x, cube.origin.y, cube.origin.z};
    orc_tree* tree = orc_tree_build(n, sx.data(), sy.data(), sz.data(), sm.data(), codes.data(),
                                    origin, cube.size, theta, std::max<uint64_t>(cap, 1), 0);
    const V3<double> c0 = cube.centre();
    const std::size_t M = tree->nodes.size();
    std::vector<float> nx(M), ny(M), nz(M), acc(M);
    for (std::size_t k = 0; k < M; ++k) {
        nx[k] = static_cast<float>(tree->nodes[k].com.x - c0.x);
        ny[k] = static_cast<float>(tree->nodes[k].com.y - c0.y);
        nz[k] = static_cast<float>(tree->nodes[k].com.z - c0.z);
        acc[k] = static_cast<float>(tree->nodes[k].acc_r2);
    }
    std::atomic<std::uint64_t> pp{0}, pc{0};
    Executor ex(resolve_threads(threads));
    ex.run(n, [&](std::size_t b, std::size_t e) {
        std::uint64_t lpp = 0, lpc = 0;
        for (std::size_t i = b; i < e; ++i) {
            const float xi = static_cast<float>(sx[i] - c0.x);
            const float yi = static_cast<float>(sy[i] - c0.y);
            const float zi = static_cast<float>(sz[i] - c0.z);
            std::size_t node = 0;
            while (node < M) {
                const Node& c = tree->nodes[node];
                const float dx = nx[node] - xi, dy = ny[node] - yi, dz = nz[node] - zi;
                const float r2 = std::fmaf(dz, dz, std::fmaf(dy, dy, dx * dx));
                if (r2 >= acc[node]) {
                    ++lpc;
                    node = c.next;
                } else if (c.count > 0) {
                    lpp += c.count - ((i >= c.first && i < c.first + c.count) ? 1 : 0);
                    node = c.next;
                } else {
                    ++node;
                }
            }
        }
        pp.fetch_add(lpp);
        pc.fetch_add(lpc);
    });
    orc_tree_free(tree);
    counts[0] = pp.load();
    counts[1] = pc.load();
    return 0;
}

int orc_refresh(uint64_t n, const double* x, const double* y, const double* z, const double* m,
                const orc_field* f, double* ax, double* ay, double* az) {
    if (f == nullptr) return 1;
    field_eval(n, x, y, z, m, *f, ax, ay, az);
    return 0;
}

int orc_steps(uint64_t n, double* x, double* y, double* z, double* vx, double* vy, double* vz,
              double* ax, double* ay, double* az, const double* m, const orc_field* f,
              double dt, uint64_t steps) {
    if (f == nullptr) return 1;
    for (uint64_t s = 0; s < steps; ++s) {
        if (f->integrator == 2) {
            rk4_step(n, x, y, z, vx, vy, vz, ax, ay, az, m, *f, dt);
        } else if (f->integrator == 1) {
            vv_step(n, x, y, z, vx, vy, vz, ax, ay, az, m, *f, kOuter * dt);
            vv_step(n, x, y, z, vx, vy, vz, ax, ay, az, m, *f, kInner * dt);
            vv_step(n, x, y, z, vx, vy, vz, ax, ay, az, m, *f, kOuter * dt);
        } else {
            vv_step(n, x, y, z, vx, vy, vz, ax, ay, az, m, *f, dt);
        }
    }
    return 0;
}

int orc_diagnostics(uint64_t n, const double* x, const double* y, const double* z,
                    const double* vx, const double* vy, const double* vz, const double* m,
                    double e2, double G, double out[8]) {
    CompensatedSum ke, px, py, pz, lx, ly, lz;
    for (uint64_t i = 0; i < n; ++i) {
        ke.add(m[i] * ((vx[i] * vx[i]) + (vy[i] * vy[i]) + (vz[i] * vz[i])));
        px.add(vx[i] * m[i]);
        py.add(vy[i] * m[i]);
        pz.add(vz[i] * m[i]);
        // cross(position, velocity) * mass (vec3.hpp cross).
        lx.add(((y[i] * vz[i]) - (z[i] * vy[i])) * m[i]);
        ly.add(((z[i] * vx[i]) - (x[i] * vz[i])) * m[i]);
        lz.add(((x[i] * vy[i]) - (y[i] * vx[i])) * m[i]);
    }
    // diagnostics.cpp:129-158: rows i, inner plain sum over j > i.
    CompensatedSum pe;
    for (uint64_t i = 0; i < n; ++i) {
        double inter = 0;
        for (uint64_t j = i + 1; j < n; ++j) {
            const double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            inter += m[j] * (1.0 / std::sqrt(((dx * dx) + (dy * dy) + (dz * dz)) + e2));
        }
        pe.add(-G * m[i] * inter);
    }
    out[0] = ke.total() / 2;
    out[1] = pe.total();
    out[2] = px.total();
    out[3] = py.total();
    out[4] = pz.total();
    out[5] = lx.total();
    out[6] = ly.total();
    out[7] = lz.total();
    return 0;
}

}  // extern "C"
