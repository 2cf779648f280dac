#pragma once
// Minimal restatement of the reference's Softening (core/softening.hpp:50-78):
// holds the square of the softening length.
#include <cmath>

#include "orrery/core/types.hpp"

namespace orrery::core {

class Softening {
public:
    constexpr Softening() noexcept = default;
    constexpr explicit Softening(Real length) noexcept : squared_(length * length) {}
    [[nodiscard]] constexpr Real squared() const noexcept { return squared_; }
    [[nodiscard]] Real length() const noexcept { return std::sqrt(squared_); }

private:
    Real squared_{};
};

}  // namespace orrery::core
