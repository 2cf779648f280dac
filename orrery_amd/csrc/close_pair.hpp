// Close pairs in FP64.
//
// Where two particles are much closer than their tiles are wide, their one
// term can be many times the net force on either particle (a bound pair near
// the centre of a cluster). An FP32 term carries a relative error of a few
// 1e-8 from its own roundings, and that error, scaled by the term over the net
// force, is what the worst particles' error is made of. These pairs are
// evaluated in FP64 instead, from the positions the kernels hold as (hi, lo)
// float pairs relative to one tile centre:
//
//   D  = (b_hi + b_lo) - (a_hi + a_lo)   each sum exact in FP64, one rounding,
//   R  = |D|^2 + eps^2                   by FMA,
//   y  = 1 / sqrt(R)                     the hardware FP64 reciprocal square
//                                        root and two Newton steps,
//   G  = y^3 D                           the term per unit mass, about 1e-15
//                                        relative.
//
// FP64 instructions issue at the rate of unpacked FP32 ones on this GPU, so
// the term costs some 30 of them: less than an exact double-float FP32
// evaluation of the same term, and more accurate. The caller scales G by the
// source's (or target's) mass and adds it to its FP32 sums as (hi, lo).
//
// Host and device: the host build seeds the reciprocal square root with
// 1/sqrt(x) (a CPU test checks the term against long-double arithmetic).
#pragma once

#include <cmath>

#ifndef __HIPCC__
#define ORR_HD
#else
#define ORR_HD __host__ __device__
#endif

namespace orr {

ORR_HD inline double seed_rsq64(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsq(x);
#else
    return 1.0 / std::sqrt(x);
#endif
}

// G = (|D|^2 + eps2)^(-3/2) D for the separation D = b - a of two positions
// given as (hi, lo) float pairs in one frame. R > 0 is the caller's condition
// (a close pair is never a coincident, unsoftened one).
ORR_HD inline void close_pair(const float ah[3], const float al[3], const float bh[3],
                              const float bl[3], float eps2, double G[3]) {
#pragma clang fp contract(off)
    double d[3];
    for (int c = 0; c < 3; ++c) {
        d[c] = (static_cast<double>(bh[c]) + static_cast<double>(bl[c])) -
               (static_cast<double>(ah[c]) + static_cast<double>(al[c]));
    }
    const double r2 = __builtin_fma(d[2], d[2], __builtin_fma(d[1], d[1], __builtin_fma(d[0], d[0], static_cast<double>(eps2))));
    const double h = 0.5 * r2;
    double y = seed_rsq64(r2);
    for (int k = 0; k < 2; ++k) {
        const double e = __builtin_fma(-h, y * y, 0.5);
        y = __builtin_fma(y, e, y);
    }
    const double g = (y * y) * y;
    for (int c = 0; c < 3; ++c) G[c] = g * d[c];
}

// v split into floats hi + lo (hi the rounding of v, lo that of the rest).
ORR_HD inline void split(double v, float& hi, float& lo) {
    hi = static_cast<float>(v);
    lo = static_cast<float>(v - static_cast<double>(hi));
}

}  // namespace orr
