#pragma once
/// \file
/// The engine's two `[solver] kind` values and the assembly step that turns one
/// into a ForceSolver.
///
/// The reference's run assembly picks a GPU solver in `try_make_gpu_solver`
/// (src/sim/assembly.cpp:116-129), decides what a null means in `make_solver`
/// (assembly.cpp:246-272: refuse when `allow_cpu_fallback` is false, otherwise
/// print a note and build the CPU equivalent of `cpu_equivalent`,
/// assembly.cpp:84-91), and spells the kinds in `to_string(SolverKind)`
/// (src/sim/configuration.cpp:101-112). These functions are that logic for
/// `hip-direct` and `hip-tree`, so the reference's assembly can call them
/// without knowing the engine: the CPU solver itself stays the caller's to build.
#include <memory>
#include <optional>
#include <ostream>
#include <string_view>

#include "orrery/core/softening.hpp"
#include "orrery/solvers/force_solver.hpp"
#include "orrery/solvers/octree.hpp"

namespace orrery::solvers {

/// The engine's solver kinds, beside the reference's `SolverKind`
/// (include/orrery/sim/configuration.hpp:56-61).
enum class HipSolverKind : unsigned char {
    kHipDirect,
    kHipTree,
};

/// "hip-direct" / "hip-tree": the `[solver] kind` spellings.
[[nodiscard]] std::string_view to_string(HipSolverKind kind) noexcept;

/// The kind whose spelling is `text` (exact, case-sensitive, as
/// configuration.cpp's `parse_kind`), or nothing.
[[nodiscard]] std::optional<HipSolverKind> parse_hip_solver_kind(std::string_view text) noexcept;

/// Whether the kind walks a tree (configuration.cpp:31-33 `is_tree_solver`),
/// i.e. reads `opening_angle`, `leaf_capacity` and `quadrupole`.
[[nodiscard]] bool is_tree_solver(HipSolverKind kind) noexcept;

/// The CPU solver computing the same physics: "barnes-hut" for the tree,
/// "direct" otherwise (assembly.cpp:84-91).
[[nodiscard]] std::string_view cpu_equivalent(HipSolverKind kind) noexcept;

/// The `[solver]` settings the engine reads (SolverSettings,
/// configuration.hpp:236-273, minus the CPU executor's).
struct HipSolverSettings {
    HipSolverKind kind = HipSolverKind::kHipTree;
    double softening = 0.0;
    double opening_angle = 0.5;
    unsigned long long leaf_capacity = 32;
    bool quadrupole = false;
    bool allow_cpu_fallback = true;
    int device = 0;
};

/// The GPU solver, or null when this machine has no usable device
/// (`try_make_gpu_solver`, assembly.cpp:116-129). Tree parameters are corrected
/// as the reference corrects them (octree.cpp:189-194).
[[nodiscard]] std::unique_ptr<ForceSolver> try_make_hip_solver(const HipSolverSettings& settings);

/// `make_solver`'s GPU branch (assembly.cpp:246-272): the solver if there is a
/// device; otherwise a std::runtime_error when `allow_cpu_fallback` is false,
/// or a null after writing the reference's one-line note to `report`, in which
/// case the caller builds `cpu_equivalent(settings.kind)` on the CPU.
[[nodiscard]] std::unique_ptr<ForceSolver> make_hip_solver(const HipSolverSettings& settings,
                                                           std::ostream& report);

}  // namespace orrery::solvers
