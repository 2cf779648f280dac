#pragma once
// Minimal restatement of the reference's Vec3 (core/vec3.hpp:27-40) and
// Vec3Span (core/vec3_span.hpp:33-74): three equal-length component spans.
#include <span>

#include "orrery/core/types.hpp"

namespace orrery::core {

struct Vec3 {
    Real x{};
    Real y{};
    Real z{};
};

template<typename T> struct Vec3Span {
    std::span<T> x;
    std::span<T> y;
    std::span<T> z;
    [[nodiscard]] constexpr Index size() const noexcept { return x.size(); }
    [[nodiscard]] constexpr bool empty() const noexcept { return x.empty(); }
    constexpr operator Vec3Span<const T>() const noexcept
        requires(!std::is_const_v<T>)
    {
        return {x, y, z};
    }
};

}  // namespace orrery::core
