#pragma once
// Minimal restatement of the reference's ForceSolver
// (solvers/force_solver.hpp:41-87) and InteractionCount
// (solvers/interaction_count.hpp:37-76).
#include <cstdint>
#include <string_view>

#include "orrery/core/softening.hpp"
#include "orrery/integrators/acceleration_field.hpp"
#include "orrery/solvers/interaction_count.hpp"

namespace orrery::solvers {

class ForceSolver : public integrators::AccelerationField {
public:
    ~ForceSolver() override = default;
    [[nodiscard]] virtual std::string_view name() const noexcept = 0;
    [[nodiscard]] virtual core::Softening softening() const noexcept = 0;
    [[nodiscard]] virtual InteractionCount interaction_count() const noexcept = 0;
    virtual void reset_interaction_count() noexcept = 0;

protected:
    ForceSolver() = default;
    ForceSolver(const ForceSolver&) = default;
    ForceSolver(ForceSolver&&) = default;
    ForceSolver& operator=(const ForceSolver&) = default;
    ForceSolver& operator=(ForceSolver&&) = default;
};

}  // namespace orrery::solvers
