#pragma once
// The reference's unit choice (core/units.hpp:38): G = 1.
#include "orrery/core/types.hpp"

namespace orrery::core {
inline constexpr Real kGravitationalConstant = 1;
}  // namespace orrery::core
