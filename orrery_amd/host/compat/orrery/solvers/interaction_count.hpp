#pragma once
// Minimal restatement of the reference's InteractionCount (solvers/interaction_count.hpp:37-76).
#include <cstdint>

namespace orrery::solvers {

struct InteractionCount {
    std::uint64_t evaluations{};
    std::uint64_t particle_particle{};
    std::uint64_t particle_cell{};
};

}  // namespace orrery::solvers
