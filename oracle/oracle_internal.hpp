// Internal declarations shared by the oracle's translation units.
// Test infrastructure only: see oracle.h.
#pragma once

#include <cstddef>
#include <cstdint>

namespace oracle {

template<typename R> struct V3 {
    R x{};
    R y{};
    R z{};
};

// Source arrays of one direct sum (structure of arrays).
template<typename R> struct Sources {
    const R* x;
    const R* y;
    const R* z;
    const R* m;
};

// direct_kernel_scalar.cpp:31-59: sum over [begin, end) in index order, no G.
template<typename R>
V3<R> accumulate_scalar(const Sources<R>& s, V3<R> target, std::size_t begin, std::size_t end,
                        R softening_squared) noexcept;

// direct_kernel_avx2.cpp:169-230: 4 FP64 / 8 FP32 lanes, FMA accumulate, fixed
// lane fold, scalar remainder. Lives in the -mavx2 -mfma translation unit.
V3<double> accumulate_avx2(const Sources<double>& s, V3<double> target, std::size_t begin,
                           std::size_t end, double softening_squared) noexcept;
V3<float> accumulate_avx2(const Sources<float>& s, V3<float> target, std::size_t begin,
                          std::size_t end, float softening_squared) noexcept;

}  // namespace oracle
