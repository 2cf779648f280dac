// Close pairs in double-float FP32 arithmetic.
//
// Where two particles are much closer than their tiles are wide, their one
// term can be many times the net force on either particle (a bound pair near
// the centre of a cluster). An FP32 term carries a relative error of a few
// 1e-8 from rounding the separation alone, and that error, scaled by the term
// over the net force, is what the worst particles' error is made of. These
// helpers evaluate such a term with every quantity held as an unevaluated sum
// hi + lo of two floats (Dekker's and Knuth's error-free transformations:
// TwoSum, TwoProd by FMA), to about 1e-12 relative:
//
//   D  = b - a                        from the two positions' (hi, lo) parts,
//   R  = |D|^2 + eps^2                summed with TwoSum, squares by TwoProd,
//   y0 = rsq(R_hi)                    the hardware reciprocal square root,
//   e  = 1 - R y0^2                   the residual, computed exactly enough,
//   G  = y0^3 (1 + 3e/2)              one Newton step on y0^3 (O(e^2) left),
//   T  = m G D                        products by TwoProd.
//
// Host and device: the host build uses 1/sqrt(x) for the seed (a CPU test
// checks the transformation against long-double arithmetic).
#pragma once

#include <cmath>

#ifndef __HIPCC__
#define ORR_HD
#define ORR_FMAF(a, b, c) std::fma(a, b, c)
#else
#define ORR_HD __host__ __device__
#define ORR_FMAF(a, b, c) __builtin_fmaf(a, b, c)
#endif

namespace orr {

struct DF {
    float hi, lo;
};

ORR_HD inline void two_sum(float a, float b, float& s, float& e) {
#pragma clang fp contract(off)
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

ORR_HD inline void two_prod(float a, float b, float& p, float& e) {
#pragma clang fp contract(off)
    p = a * b;
    e = ORR_FMAF(a, b, -p);
}

ORR_HD inline float seed_rsq(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / std::sqrt(x);
#endif
}

// Separation D = b - a of two positions given as (hi, lo) float pairs in one
// frame, and G = (|D|^2 + eps2)^(-3/2), both as double-floats.
ORR_HD inline void close_pair(const float ah[3], const float al[3], const float bh[3],
                              const float bl[3], float eps2, DF D[3], DF& G) {
#pragma clang fp contract(off)
    float rh = eps2, rl = 0.f;
    for (int c = 0; c < 3; ++c) {
        // (bh + bl) - (ah + al) as a double-float: both differences exact with
        // their error terms, then renormalised twice. For the closest pairs the
        // lo parts' difference is as large as the hi parts', so it is not
        // rounded on its own.
        float s, e, t, f;
        two_sum(bh[c], -ah[c], s, e);
        two_sum(bl[c], -al[c], t, f);
        two_sum(s, e + t, s, e);
        two_sum(s, e + f, D[c].hi, D[c].lo);
        float p, pe;
        two_prod(D[c].hi, D[c].hi, p, pe);
        float q, qe;
        two_sum(rh, p, q, qe);
        rh = q;
        rl = rl + (qe + (pe + 2.f * D[c].hi * D[c].lo));
    }
    {
        const float s = rh + rl;
        rl = rl - (s - rh);
        rh = s;
    }
    const float y0 = seed_rsq(rh);
    float y2, y2e;
    two_prod(y0, y0, y2, y2e);
    float p, pe;
    two_prod(rh, y2, p, pe);
    // p is within a few ulps of 1, so 1 - p is exact.
    const float e = (1.f - p) - (pe + (rh * y2e + rl * y2));
    float c3, c3e;
    two_prod(y2, y0, c3, c3e);
    c3e = ORR_FMAF(y2e, y0, c3e);
    G.hi = c3;
    G.lo = c3e + c3 * (1.5f * e);
}

// T = m G D as double-floats (m exact in FP32).
ORR_HD inline void close_term(float m, const DF& G, const DF D[3], DF T[3]) {
#pragma clang fp contract(off)
    float fh, fe;
    two_prod(m, G.hi, fh, fe);
    const float fl = fe + m * G.lo;
    for (int c = 0; c < 3; ++c) {
        float th, te;
        two_prod(fh, D[c].hi, th, te);
        T[c].hi = th;
        T[c].lo = te + (fh * D[c].lo + fl * D[c].hi);
    }
}

}  // namespace orr
