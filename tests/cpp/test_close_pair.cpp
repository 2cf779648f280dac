// The FP64 close-pair term of the direct kernels (orrery_amd/csrc/close_pair.hpp)
// against long-double arithmetic: separations from 1e-7 to 1 of the positions'
// magnitude, softenings from 0 to larger than the separation.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../orrery_amd/csrc/close_pair.hpp"

int main() {
    std::mt19937_64 rng(20260811);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::uniform_real_distribution<double> lg(-7.0, 0.0);
    double worst = 0, worst_fp32 = 0;
    int count = 0;
    for (int k = 0; k < 200000; ++k) {
        const double scale = std::pow(10.0, lg(rng) * 0.3);  // positions' magnitude
        double a[3], b[3];
        const double sep = scale * std::pow(10.0, lg(rng));
        for (int c = 0; c < 3; ++c) {
            a[c] = scale * u(rng);
            b[c] = a[c] + sep * u(rng);
        }
        const float eps = static_cast<float>(k % 4 == 0 ? 0.0 : sep * std::pow(10.0, 2.0 * u(rng)));
        const float eps2 = eps * eps;
        float ah[3], al[3], bh[3], bl[3];
        for (int c = 0; c < 3; ++c) {
            ah[c] = static_cast<float>(a[c]);
            al[c] = static_cast<float>(a[c] - ah[c]);
            bh[c] = static_cast<float>(b[c]);
            bl[c] = static_cast<float>(b[c] - bh[c]);
        }
        // The exact term for the positions the pairs represent.
        long double d[3], r2 = static_cast<long double>(eps2);
        for (int c = 0; c < 3; ++c) {
            d[c] = (static_cast<long double>(bh[c]) + bl[c]) - (static_cast<long double>(ah[c]) + al[c]);
            r2 += d[c] * d[c];
        }
        if (!(r2 > 0)) continue;
        const double m = 0.75;
        const long double g = 1.0L / (r2 * std::sqrt(r2));
        double G[3];
        orr::close_pair(ah, al, bh, bl, eps2, G);
        long double err = 0, mag = 0, err32 = 0;
        for (int c = 0; c < 3; ++c) {
            const long double exact = m * g * d[c];
            // As the kernels add it: m G split into two floats.
            float hi, lo;
            orr::split(m * G[c], hi, lo);
            const long double got = static_cast<long double>(hi) + lo;
            err += (got - exact) * (got - exact);
            mag += exact * exact;
        }
        {
            // The plain FP32 term on the rounded separation, for comparison.
            float dd[3];
            for (int c = 0; c < 3; ++c) dd[c] = (bh[c] - ah[c]) + (bl[c] - al[c]);
            float rr = std::fma(dd[0], dd[0], eps2);
            rr = std::fma(dd[1], dd[1], rr);
            rr = std::fma(dd[2], dd[2], rr);
            const float inv = 1.0f / std::sqrt(rr);
            const float f = (static_cast<float>(m) * inv) * (inv * inv);
            for (int c = 0; c < 3; ++c) {
                const long double e = static_cast<long double>(f * dd[c]) - m * g * d[c];
                err32 += e * e;
            }
        }
        if (mag == 0) continue;
        const double rel = static_cast<double>(std::sqrt(err / mag));
        const double rel32 = static_cast<double>(std::sqrt(err32 / mag));
        if (rel > worst) worst = rel;
        if (rel32 > worst_fp32) worst_fp32 = rel32;
        ++count;
    }
    std::printf("close pairs: %d cases, worst relative error %.3e (plain FP32 term %.3e)\n", count,
                worst, worst_fp32);
    return worst <= 1e-11 && count > 100000 ? 0 : 1;
}
