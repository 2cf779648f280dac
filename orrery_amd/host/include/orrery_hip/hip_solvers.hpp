#pragma once
/// \file
/// MI355X force solvers behind the reference's ForceSolver seam.
///
/// `HipDirectSolver` and `HipTreeSolver` are drop-ins for the reference's
/// CudaDirectSolver / CudaTreeSolver (include/orrery/solvers/cuda_direct_solver.hpp,
/// cuda_tree_solver.hpp): the same virtuals, the same `try_create` null-on-no-device
/// factory (cuda_direct_solver.cpp:140-164), the same write-not-accumulate and
/// synchronous-evaluation contract (acceleration_field.hpp:48-49;
/// cuda_kernels.hpp:108-110). The device work is the C ABI of include/orrery_hip.h;
/// failures surface as `HipError` (the reference's backend::CudaError,
/// cuda_memory.hpp:129-153), counts beyond 2^31 as std::length_error.
///
/// Compiled against the reference's own headers inside its build, or against
/// orrery_amd/host/compat (minimal restatements of those headers) standalone.
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <string_view>

#include "orrery/core/softening.hpp"
#include "orrery/core/types.hpp"
#include "orrery/core/vec3_span.hpp"
#include "orrery/solvers/force_solver.hpp"
#include "orrery/solvers/interaction_count.hpp"
#include "orrery/solvers/octree.hpp"

namespace orrery::solvers {

/// A failed engine call: the status of include/orrery_hip.h and its message.
class HipError : public std::runtime_error {
public:
    HipError(int status, const std::string& message)
        : std::runtime_error(message), status_(status) {}
    [[nodiscard]] int status() const noexcept { return status_; }

private:
    int status_;
};

/// Per-phase device time of the evaluations since the last reset, in
/// milliseconds, measured with hipEvents on the engine's stream while timing
/// is on (`set_timing(true)`). Timing is off by default: then repeated
/// evaluations of one size replay a captured hipGraph (counted in
/// `graph_launches`) and the phase times stay zero; with timing on every
/// evaluation runs eagerly so that its phases can be timed. The results are
/// the same bits either way.
struct HipTimings {
    double prepare{};
    double kernel{};
    double reduce{};
    double ordering{};     // bounding cube, Morton encode and sort
    double construction{}; // tree build and moments
    double gathering{};
    double traversal{};
    std::uint64_t graph_launches{};
};

/// A summary of the last tree (Octree::leaf_count / depth, octree.hpp:229-236).
struct HipTreeSummary {
    std::uint64_t node_count{};
    std::uint64_t leaf_count{};
    unsigned depth{};
};

/// Direct O(N^2) summation on the GPU; `name()` is "hip-direct".
class HipDirectSolver final : public ForceSolver {
public:
    [[nodiscard]] static std::unique_ptr<HipDirectSolver> try_create(core::Softening softening = {},
                                                                     int device = 0);
    ~HipDirectSolver() override;
    HipDirectSolver(const HipDirectSolver&) = delete;
    HipDirectSolver& operator=(const HipDirectSolver&) = delete;
    HipDirectSolver(HipDirectSolver&&) noexcept;
    HipDirectSolver& operator=(HipDirectSolver&&) noexcept;

    void evaluate(core::Vec3Span<const core::Real> positions, std::span<const core::Real> masses,
                  core::Vec3Span<core::Real> accelerations) override;
    [[nodiscard]] std::string_view name() const noexcept override { return "hip-direct"; }
    [[nodiscard]] core::Softening softening() const noexcept override;
    [[nodiscard]] InteractionCount interaction_count() const noexcept override;
    void reset_interaction_count() noexcept override;

    [[nodiscard]] HipTimings timings() const;
    void reset_timings();
    void set_timing(bool enabled);

private:
    struct Impl;
    explicit HipDirectSolver(std::unique_ptr<Impl> impl) noexcept;
    std::unique_ptr<Impl> impl_;
};

/// Barnes-Hut with on-device ordering, build and walk; `name()` is "hip-tree".
class HipTreeSolver final : public ForceSolver {
public:
    [[nodiscard]] static std::unique_ptr<HipTreeSolver> try_create(TreeParameters parameters = {},
                                                                   core::Softening softening = {},
                                                                   int device = 0);
    ~HipTreeSolver() override;
    HipTreeSolver(const HipTreeSolver&) = delete;
    HipTreeSolver& operator=(const HipTreeSolver&) = delete;
    HipTreeSolver(HipTreeSolver&&) noexcept;
    HipTreeSolver& operator=(HipTreeSolver&&) noexcept;

    void evaluate(core::Vec3Span<const core::Real> positions, std::span<const core::Real> masses,
                  core::Vec3Span<core::Real> accelerations) override;
    [[nodiscard]] std::string_view name() const noexcept override { return "hip-tree"; }
    [[nodiscard]] core::Softening softening() const noexcept override;
    [[nodiscard]] InteractionCount interaction_count() const noexcept override;
    void reset_interaction_count() noexcept override;

    [[nodiscard]] const TreeParameters& parameters() const noexcept;
    [[nodiscard]] HipTreeSummary tree() const noexcept;
    [[nodiscard]] std::uint64_t node_visits() const noexcept;
    [[nodiscard]] HipTimings timings() const;
    void reset_timings();
    void set_timing(bool enabled);

private:
    struct Impl;
    explicit HipTreeSolver(std::unique_ptr<Impl> impl) noexcept;
    std::unique_ptr<Impl> impl_;
};

}  // namespace orrery::solvers
