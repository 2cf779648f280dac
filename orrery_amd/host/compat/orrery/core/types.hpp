#pragma once
// Minimal restatement of the reference's include/orrery/core/types.hpp:19-45
// (the names and meanings the solver seam uses), so the host classes in
// orrery_amd/host compile without the reference checkout. Inside the reference
// build its own header is used instead.
#include <cstddef>
#include <type_traits>

namespace orrery::core {
#ifdef ORRERY_SINGLE_PRECISION
using Real = float;
#else
using Real = double;
#endif
inline constexpr bool kSinglePrecision = std::is_same_v<Real, float>;
using Index = std::size_t;
}  // namespace orrery::core
