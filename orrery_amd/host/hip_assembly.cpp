// Assembly of the engine's solvers from `[solver]` settings (see
// include/orrery_hip/hip_assembly.hpp). The refusal and the fallback note are
// the reference's wording (src/sim/assembly.cpp:256-271) with this engine's
// kinds, so a run's terminal output reads the same whichever GPU backend it
// asked for.
#include "orrery_hip/hip_assembly.hpp"

#include <array>
#include <stdexcept>
#include <string>

#include "orrery/core/types.hpp"
#include "orrery_hip/hip_solvers.hpp"

namespace orrery::solvers {
namespace {

constexpr std::array kHipSolverKinds{HipSolverKind::kHipDirect, HipSolverKind::kHipTree};

// The engine is linked in whenever this file is, so the only reason a solver
// can be missing is the machine (the reference's other reason, a build without
// the backend, cannot arise here).
constexpr std::string_view kNoDevice = "no usable device was found";

}  // namespace

std::string_view to_string(HipSolverKind kind) noexcept {
    switch (kind) {
    case HipSolverKind::kHipDirect:
        return "hip-direct";
    case HipSolverKind::kHipTree:
        return "hip-tree";
    }
    return "hip-direct";
}

std::optional<HipSolverKind> parse_hip_solver_kind(std::string_view text) noexcept {
    for (const HipSolverKind kind : kHipSolverKinds) {
        if (to_string(kind) == text) return kind;
    }
    return std::nullopt;
}

bool is_tree_solver(HipSolverKind kind) noexcept { return kind == HipSolverKind::kHipTree; }

std::string_view cpu_equivalent(HipSolverKind kind) noexcept {
    return kind == HipSolverKind::kHipTree ? "barnes-hut" : "direct";
}

std::unique_ptr<ForceSolver> try_make_hip_solver(const HipSolverSettings& settings) {
    const core::Softening softening{static_cast<core::Real>(settings.softening)};
    if (settings.kind == HipSolverKind::kHipTree) {
        TreeParameters parameters{};
        parameters.opening_angle = static_cast<core::Real>(settings.opening_angle);
        parameters.leaf_capacity = static_cast<core::Index>(settings.leaf_capacity);
        parameters.quadrupole = settings.quadrupole;
        return HipTreeSolver::try_create(parameters, softening, settings.device);
    }
    return HipDirectSolver::try_create(softening, settings.device);
}

std::unique_ptr<ForceSolver> make_hip_solver(const HipSolverSettings& settings, std::ostream& report) {
    if (std::unique_ptr<ForceSolver> solver = try_make_hip_solver(settings)) return solver;
    const std::string kind{to_string(settings.kind)};
    if (!settings.allow_cpu_fallback) {
        throw std::runtime_error("solver.kind is " + kind + " and " + std::string{kNoDevice} +
                                 ". Set solver.allow_cpu_fallback = true to run on the CPU instead");
    }
    report << "note: " << kind << " was asked for but " << kNoDevice << ", so this run uses "
           << cpu_equivalent(settings.kind) << " on the CPU\n";
    return nullptr;
}

}  // namespace orrery::solvers
