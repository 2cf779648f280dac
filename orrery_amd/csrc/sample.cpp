// The reference's seeded Plummer sampler, on the host, so that a device run is
// built from exactly the particles the reference's CPU path would draw for the
// same seed (src/initial_conditions/plummer.cpp:91-128 over the random source
// of src/core/random.cpp:12-47, then centre_of_mass_frame.cpp).
//
// Host C++ only (compiled by the host compiler with contraction off and the
// sine and cosine left as separate libm calls), because bit equality with the
// reference sample rests on the same libm calls in the same order:
//
//   * the engine is std::mt19937_64 seeded with the seed itself; a uniform
//     draw keeps the top 53 bits and scales them by 2^-53; uniform(lo, hi) is
//     the interpolation lo (1 - f) + hi f;
//   * per particle: the mass fraction uniform(0, cutoff); the radius
//     a / sqrt(pow(f, -2/3) - 1); speed-fraction pairs (q, envelope draw) until
//     0.1 u < q^2 c^3 sqrt(c) with c = 1 - q^2; then the position direction and
//     after it the velocity direction (the order Clang gives the two draws
//     inside one argument list at plummer.cpp:122), each a height uniform(-1, 1)
//     then an azimuth uniform(0, 2 pi);
//   * the centre-of-mass shift uses Neumaier-compensated sums of m, m x and
//     m v in index order (core/diagnostics.cpp centre_of_mass /
//     linear_momentum).

#include <cmath>
#include <cstdint>
#include <numbers>
#include <random>

#include "../../include/orrery_hip.h"

#pragma STDC FP_CONTRACT OFF

namespace orr {
void set_last_error(const char* what);
}

namespace {

class Draws {
public:
    explicit Draws(std::uint64_t seed) : engine_(seed) {}

    double unit() {
        constexpr double kScale = 1.0 / 9007199254740992.0;  // 2^-53
        return static_cast<double>(engine_() >> 11) * kScale;
    }
    double between(double lo, double hi) {
        const double f = unit();
        return (lo * (1.0 - f)) + (hi * f);
    }
    void direction(double& ux, double& uy, double& uz) {
        const double height = between(-1.0, 1.0);
        const double azimuth = between(0.0, 2.0 * std::numbers::pi_v<double>);
        const double ring = std::sqrt(1.0 - (height * height));
        ux = ring * std::cos(azimuth);
        uy = ring * std::sin(azimuth);
        uz = height;
    }

private:
    std::mt19937_64 engine_;
};

// Neumaier's compensated running sum.
struct Sum {
    double total = 0.0, carry = 0.0;
    void add(double v) {
        const double t = total + v;
        carry += std::fabs(total) >= std::fabs(v) ? (total - t) + v : (v - t) + total;
        total = t;
    }
    double value() const { return total + carry; }
};

double shift_of(const double* v, const double* m, std::uint64_t n, double mass) {
    Sum s;
    for (std::uint64_t i = 0; i < n; ++i) s.add(v[i] * m[i]);
    return mass != 0.0 ? s.value() / mass : 0.0;
}

}  // namespace

extern "C" ORR_API int orr_plummer_sample(uint64_t count, double total_mass, double scale_radius,
                                          double mass_fraction_cutoff, uint64_t seed, double* x,
                                          double* y, double* z, double* vx, double* vy, double* vz,
                                          double* m) {
    if (!(total_mass > 0) || !(scale_radius > 0) || !(mass_fraction_cutoff > 0) ||
        !(mass_fraction_cutoff < 1)) {
        orr::set_last_error("orr_plummer_sample: needs a positive mass and scale radius and a "
                            "mass fraction cutoff in (0, 1)");
        return ORR_ERR_INVALID_ARGUMENT;
    }
    if (count == 0) return ORR_OK;
    if (x == nullptr || y == nullptr || z == nullptr || vx == nullptr || vy == nullptr ||
        vz == nullptr || m == nullptr) {
        orr::set_last_error("orr_plummer_sample: null output array");
        return ORR_ERR_INVALID_ARGUMENT;
    }
    Draws draw(seed);
    const double a = scale_radius;
    const double pm = total_mass / static_cast<double>(count);
    for (std::uint64_t k = 0; k < count; ++k) {
        const double f = draw.between(0.0, mass_fraction_cutoff);
        const double r = a / std::sqrt(std::pow(f, -2.0 / 3.0) - 1.0);
        double q = 0.0;
        for (;;) {
            q = draw.unit();
            const double c = 1.0 - (q * q);
            const double density = q * q * c * c * c * std::sqrt(c);
            if (0.1 * draw.unit() < density) break;
        }
        const double escape = std::sqrt(2.0 * 1.0 * total_mass / std::sqrt((r * r) + (a * a)));
        const double speed = q * escape;
        double px, py, pz, qx, qy, qz;
        draw.direction(px, py, pz);
        draw.direction(qx, qy, qz);
        x[k] = px * r;
        y[k] = py * r;
        z[k] = pz * r;
        vx[k] = qx * speed;
        vy[k] = qy * speed;
        vz[k] = qz * speed;
        m[k] = pm;
    }
    Sum mass;
    for (std::uint64_t i = 0; i < count; ++i) mass.add(m[i]);
    const double total = mass.value();
    const double cx = shift_of(x, m, count, total), cy = shift_of(y, m, count, total),
                 cz = shift_of(z, m, count, total);
    const double dx = shift_of(vx, m, count, total), dy = shift_of(vy, m, count, total),
                 dz = shift_of(vz, m, count, total);
    for (std::uint64_t i = 0; i < count; ++i) {
        x[i] = x[i] - cx;
        y[i] = y[i] - cy;
        z[i] = z[i] - cz;
        vx[i] = vx[i] - dx;
        vy[i] = vy[i] - dy;
        vz[i] = vz[i] - dz;
    }
    return ORR_OK;
}
