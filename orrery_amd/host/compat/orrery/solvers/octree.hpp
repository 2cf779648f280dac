#pragma once
// Minimal restatement of the reference's TreeParameters and corrected_parameters
// (solvers/octree.hpp:147-199; octree.cpp:189-194). The Octree class itself is
// not restated: the device engine builds its own tree.
#include <algorithm>

#include "orrery/core/types.hpp"

namespace orrery::solvers {

struct TreeParameters {
    core::Real opening_angle{static_cast<core::Real>(0.5)};
    core::Index leaf_capacity{32};
    bool quadrupole{false};
};

[[nodiscard]] inline TreeParameters corrected_parameters(TreeParameters parameters) noexcept {
    parameters.opening_angle = std::clamp(parameters.opening_angle, core::Real{0}, core::Real{1});
    parameters.leaf_capacity = std::max<core::Index>(parameters.leaf_capacity, 1);
    return parameters;
}

}  // namespace orrery::solvers
