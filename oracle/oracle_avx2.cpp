// AVX2 direct kernel of the oracle, compiled with -mavx2 -mfma.
// Test infrastructure only: see oracle.h.
//
// Follows src/solvers/direct_kernel_avx2.cpp:169-230: separations from an
// unaligned load minus the broadcast target, r^2 as dx*dx then two FMAs, plus
// eps^2, 1/sqrt, cube as (inv*inv)*inv, times mass, FMA into the accumulator.
// The lane fold (:150-161) is (v0+v2)+(v1+v3) for FP64 and the 8->4->2->1
// pairwise fold for FP32; the tail goes to the scalar kernel, which lives in
// a translation unit without -mfma so that it never contracts.
#include <immintrin.h>

#include "oracle_internal.hpp"

namespace oracle {

namespace {

double fold(__m256d v) noexcept {
    const __m128d folded = _mm_add_pd(_mm256_castpd256_pd128(v), _mm256_extractf128_pd(v, 1));
    return _mm_cvtsd_f64(_mm_add_pd(folded, _mm_unpackhi_pd(folded, folded)));
}

float fold(__m256 v) noexcept {
    __m128 folded = _mm_add_ps(_mm256_castps256_ps128(v), _mm256_extractf128_ps(v, 1));
    folded = _mm_add_ps(folded, _mm_movehl_ps(folded, folded));
    return _mm_cvtss_f32(_mm_add_ss(folded, _mm_shuffle_ps(folded, folded, _MM_SHUFFLE(1, 1, 1, 1))));
}

}  // namespace

V3<double> accumulate_avx2(const Sources<double>& s, V3<double> target, std::size_t begin,
                           std::size_t end, double softening_squared) noexcept {
    constexpr std::size_t kLanes = 4;
    const __m256d tx = _mm256_set1_pd(target.x);
    const __m256d ty = _mm256_set1_pd(target.y);
    const __m256d tz = _mm256_set1_pd(target.z);
    const __m256d e2 = _mm256_set1_pd(softening_squared);
    const __m256d one = _mm256_set1_pd(1.0);
    __m256d ax = _mm256_setzero_pd();
    __m256d ay = _mm256_setzero_pd();
    __m256d az = _mm256_setzero_pd();
    const std::size_t vector_end = begin + (((end - begin) / kLanes) * kLanes);
    for (std::size_t j = begin; j < vector_end; j += kLanes) {
        const __m256d dx = _mm256_sub_pd(_mm256_loadu_pd(s.x + j), tx);
        const __m256d dy = _mm256_sub_pd(_mm256_loadu_pd(s.y + j), ty);
        const __m256d dz = _mm256_sub_pd(_mm256_loadu_pd(s.z + j), tz);
        __m256d r2 = _mm256_mul_pd(dx, dx);
        r2 = _mm256_fmadd_pd(dy, dy, r2);
        r2 = _mm256_fmadd_pd(dz, dz, r2);
        r2 = _mm256_add_pd(r2, e2);
        const __m256d inv = _mm256_div_pd(one, _mm256_sqrt_pd(r2));
        const __m256d inv3 = _mm256_mul_pd(_mm256_mul_pd(inv, inv), inv);
        const __m256d f = _mm256_mul_pd(_mm256_loadu_pd(s.m + j), inv3);
        ax = _mm256_fmadd_pd(f, dx, ax);
        ay = _mm256_fmadd_pd(f, dy, ay);
        az = _mm256_fmadd_pd(f, dz, az);
    }
    const V3<double> rest = accumulate_scalar<double>(s, target, vector_end, end, softening_squared);
    return {fold(ax) + rest.x, fold(ay) + rest.y, fold(az) + rest.z};
}

V3<float> accumulate_avx2(const Sources<float>& s, V3<float> target, std::size_t begin,
                          std::size_t end, float softening_squared) noexcept {
    constexpr std::size_t kLanes = 8;
    const __m256 tx = _mm256_set1_ps(target.x);
    const __m256 ty = _mm256_set1_ps(target.y);
    const __m256 tz = _mm256_set1_ps(target.z);
    const __m256 e2 = _mm256_set1_ps(softening_squared);
    const __m256 one = _mm256_set1_ps(1.0f);
    __m256 ax = _mm256_setzero_ps();
    __m256 ay = _mm256_setzero_ps();
    __m256 az = _mm256_setzero_ps();
    const std::size_t vector_end = begin + (((end - begin) / kLanes) * kLanes);
    for (std::size_t j = begin; j < vector_end; j += kLanes) {
        const __m256 dx = _mm256_sub_ps(_mm256_loadu_ps(s.x + j), tx);
        const __m256 dy = _mm256_sub_ps(_mm256_loadu_ps(s.y + j), ty);
        const __m256 dz = _mm256_sub_ps(_mm256_loadu_ps(s.z + j), tz);
        __m256 r2 = _mm256_mul_ps(dx, dx);
        r2 = _mm256_fmadd_ps(dy, dy, r2);
        r2 = _mm256_fmadd_ps(dz, dz, r2);
        r2 = _mm256_add_ps(r2, e2);
        const __m256 inv = _mm256_div_ps(one, _mm256_sqrt_ps(r2));
        const __m256 inv3 = _mm256_mul_ps(_mm256_mul_ps(inv, inv), inv);
        const __m256 f = _mm256_mul_ps(_mm256_loadu_ps(s.m + j), inv3);
        ax = _mm256_fmadd_ps(f, dx, ax);
        ay = _mm256_fmadd_ps(f, dy, ay);
        az = _mm256_fmadd_ps(f, dz, az);
    }
    const V3<float> rest = accumulate_scalar<float>(s, target, vector_end, end, softening_squared);
    return {fold(ax) + rest.x, fold(ay) + rest.y, fold(az) + rest.z};
}

}  // namespace oracle
