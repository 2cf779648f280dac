// Host side of the MI355X force solvers (see include/orrery_hip/hip_solvers.hpp).
//
// Plain C++20 over the C ABI: device staging that only grows
// (cuda_direct_solver.cpp:78-137), FP64 <-> device copies, one synchronous
// engine call per evaluation, and the reference's counting conventions
// (direct_solver.cpp:75-79: N(N-1) ordered pairs per evaluation).
#include "orrery_hip/hip_solvers.hpp"

#include <array>
#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <utility>
#include <vector>

#include "orrery/core/units.hpp"
#include "orrery_hip.h"

namespace orrery::solvers {
namespace {

using core::Index;
using core::Real;

void check(int status) {
    if (status != ORR_OK) {
        const char* message = orr_last_error();
        if (status == ORR_ERR_LENGTH) throw std::length_error(message);
        throw HipError(status, message);
    }
}

// One device context and grow-only FP64 staging for x, y, z, m, ax, ay, az.
class Device {
public:
    static std::unique_ptr<Device> try_open(int device) {
        std::int32_t count = 0;
        if (orr_device_count(&count) != ORR_OK || count <= device || device < 0) return nullptr;
        orr_ctx* ctx = nullptr;
        if (orr_ctx_create(device, &ctx) != ORR_OK) return nullptr;
        return std::unique_ptr<Device>(new Device(ctx));
    }
    ~Device() {
        for (void* p : buffers_) (void)orr_free(ctx_, p);
        (void)orr_ctx_destroy(ctx_);
    }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;

    orr_ctx* ctx() const noexcept { return ctx_; }

    // Reserves staging for n particles and uploads positions and masses.
    std::array<double*, 7> stage(core::Vec3Span<const Real> positions, std::span<const Real> masses) {
        const Index n = positions.size();
        if (n > capacity_) {
            for (void*& p : buffers_) {
                if (p != nullptr) check(orr_free(ctx_, p));
                p = nullptr;
            }
            for (void*& p : buffers_) check(orr_malloc(ctx_, n * sizeof(double), &p));
            capacity_ = n;
        }
        const std::span<const Real> in[4] = {positions.x, positions.y, positions.z, masses};
        for (int k = 0; k < 4; ++k) upload(buffers_[k], in[k]);
        std::array<double*, 7> out{};
        for (int k = 0; k < 7; ++k) out[k] = static_cast<double*>(buffers_[k]);
        return out;
    }

    void download(core::Vec3Span<Real> accelerations) {
        const std::span<Real> out[3] = {accelerations.x, accelerations.y, accelerations.z};
        for (int k = 0; k < 3; ++k) {
            if constexpr (std::is_same_v<Real, double>) {
                check(orr_download(ctx_, out[k].data(), buffers_[4 + k], out[k].size() * sizeof(double)));
            } else {
                host_.resize(out[k].size());
                check(orr_download(ctx_, host_.data(), buffers_[4 + k], host_.size() * sizeof(double)));
                for (Index i = 0; i < host_.size(); ++i) out[k][i] = static_cast<Real>(host_[i]);
            }
        }
    }

    HipTimings timings() const {
        orr_timings t{};
        check(orr_ctx_timings(ctx_, &t));
        HipTimings h;
        h.prepare = t.milliseconds[0];
        h.kernel = t.milliseconds[1];
        h.reduce = t.milliseconds[2];
        h.ordering = t.milliseconds[3] + t.milliseconds[4] + t.milliseconds[5];
        h.gathering = t.milliseconds[6];
        h.construction = t.milliseconds[7] + t.milliseconds[8];
        h.traversal = t.milliseconds[9];
        h.graph_launches = t.graph_launches;
        return h;
    }
    void reset_timings() { check(orr_ctx_reset_timings(ctx_)); }
    void set_timing(bool enabled) { check(orr_ctx_set_timing(ctx_, enabled ? 1 : 0)); }

private:
    explicit Device(orr_ctx* ctx) : ctx_(ctx) {}

    void upload(void* dst, std::span<const Real> src) {
        if constexpr (std::is_same_v<Real, double>) {
            check(orr_upload(ctx_, dst, src.data(), src.size() * sizeof(double)));
        } else {
            host_.assign(src.begin(), src.end());
            check(orr_upload(ctx_, dst, host_.data(), host_.size() * sizeof(double)));
        }
    }

    orr_ctx* ctx_;
    std::array<void*, 7> buffers_{};
    Index capacity_ = 0;
    std::vector<double> host_;
};

void require_equal_lengths(core::Vec3Span<const Real> p, std::span<const Real> m,
                           core::Vec3Span<Real> a) {
    const Index n = p.size();
    if (p.y.size() != n || p.z.size() != n || m.size() != n || a.size() != n ||
        a.y.size() != n || a.z.size() != n) {
        throw std::invalid_argument("positions, masses and accelerations must have one length");
    }
    if (n > (Index{1} << 31)) throw std::length_error("more than 2^31 particles");
}

}  // namespace

// ---- HipDirectSolver ---------------------------------------------------------

struct HipDirectSolver::Impl {
    std::unique_ptr<Device> device;
    core::Softening softening;
    InteractionCount count;
};

HipDirectSolver::HipDirectSolver(std::unique_ptr<Impl> impl) noexcept : impl_(std::move(impl)) {}
HipDirectSolver::~HipDirectSolver() = default;
HipDirectSolver::HipDirectSolver(HipDirectSolver&&) noexcept = default;
HipDirectSolver& HipDirectSolver::operator=(HipDirectSolver&&) noexcept = default;

std::unique_ptr<HipDirectSolver> HipDirectSolver::try_create(core::Softening softening, int device) {
    auto dev = Device::try_open(device);
    if (dev == nullptr) return nullptr;
    auto impl = std::make_unique<Impl>();
    impl->device = std::move(dev);
    impl->softening = softening;
    return std::unique_ptr<HipDirectSolver>(new HipDirectSolver(std::move(impl)));
}

void HipDirectSolver::evaluate(core::Vec3Span<const Real> positions, std::span<const Real> masses,
                               core::Vec3Span<Real> accelerations) {
    require_equal_lengths(positions, masses, accelerations);
    const Index n = positions.size();
    if (n > 0) {
        const auto b = impl_->device->stage(positions, masses);
        orr_direct_args a{};
        a.x = b[0];
        a.y = b[1];
        a.z = b[2];
        a.m = b[3];
        a.ax = b[4];
        a.ay = b[5];
        a.az = b[6];
        a.count = n;
        a.softening_squared = static_cast<double>(impl_->softening.squared());
        a.gravitational_constant = static_cast<double>(core::kGravitationalConstant);
        a.spatial_order = 1;
        check(orr_direct_eval(impl_->device->ctx(), &a));
        impl_->device->download(accelerations);
    }
    ++impl_->count.evaluations;
    if (n > 1) {
        const auto pairs = static_cast<std::uint64_t>(n);
        impl_->count.particle_particle += pairs * (pairs - 1);
    }
}

core::Softening HipDirectSolver::softening() const noexcept { return impl_->softening; }
InteractionCount HipDirectSolver::interaction_count() const noexcept { return impl_->count; }
void HipDirectSolver::reset_interaction_count() noexcept { impl_->count = InteractionCount{}; }
HipTimings HipDirectSolver::timings() const { return impl_->device->timings(); }
void HipDirectSolver::reset_timings() { impl_->device->reset_timings(); }
void HipDirectSolver::set_timing(bool enabled) { impl_->device->set_timing(enabled); }

// ---- HipTreeSolver -----------------------------------------------------------

struct HipTreeSolver::Impl {
    std::unique_ptr<Device> device;
    TreeParameters parameters;
    core::Softening softening;
    InteractionCount count;
    std::uint64_t visits = 0;
    HipTreeSummary summary;
};

HipTreeSolver::HipTreeSolver(std::unique_ptr<Impl> impl) noexcept : impl_(std::move(impl)) {}
HipTreeSolver::~HipTreeSolver() = default;
HipTreeSolver::HipTreeSolver(HipTreeSolver&&) noexcept = default;
HipTreeSolver& HipTreeSolver::operator=(HipTreeSolver&&) noexcept = default;

std::unique_ptr<HipTreeSolver> HipTreeSolver::try_create(TreeParameters parameters,
                                                         core::Softening softening, int device) {
    auto dev = Device::try_open(device);
    if (dev == nullptr) return nullptr;
    auto impl = std::make_unique<Impl>();
    impl->device = std::move(dev);
    impl->parameters = corrected_parameters(parameters);
    impl->softening = softening;
    return std::unique_ptr<HipTreeSolver>(new HipTreeSolver(std::move(impl)));
}

void HipTreeSolver::evaluate(core::Vec3Span<const Real> positions, std::span<const Real> masses,
                             core::Vec3Span<Real> accelerations) {
    require_equal_lengths(positions, masses, accelerations);
    const Index n = positions.size();
    orr_tree_stats stats{};
    if (n > 0) {
        const auto b = impl_->device->stage(positions, masses);
        orr_tree_args a{};
        a.x = b[0];
        a.y = b[1];
        a.z = b[2];
        a.m = b[3];
        a.ax = b[4];
        a.ay = b[5];
        a.az = b[6];
        a.count = n;
        a.softening_squared = static_cast<double>(impl_->softening.squared());
        a.gravitational_constant = static_cast<double>(core::kGravitationalConstant);
        a.opening_angle = static_cast<double>(impl_->parameters.opening_angle);
        a.leaf_capacity = impl_->parameters.leaf_capacity;
        a.quadrupole = impl_->parameters.quadrupole ? 1 : 0;
        check(orr_tree_eval(impl_->device->ctx(), &a, &stats));
        impl_->device->download(accelerations);
    }
    ++impl_->count.evaluations;
    impl_->count.particle_particle += stats.particle_particle;
    impl_->count.particle_cell += stats.particle_cell;
    impl_->visits += stats.node_visits;
    impl_->summary = HipTreeSummary{stats.node_count, stats.leaf_count, stats.depth};
}

core::Softening HipTreeSolver::softening() const noexcept { return impl_->softening; }
InteractionCount HipTreeSolver::interaction_count() const noexcept { return impl_->count; }
void HipTreeSolver::reset_interaction_count() noexcept {
    impl_->count = InteractionCount{};
    impl_->visits = 0;
}
const TreeParameters& HipTreeSolver::parameters() const noexcept { return impl_->parameters; }
HipTreeSummary HipTreeSolver::tree() const noexcept { return impl_->summary; }
std::uint64_t HipTreeSolver::node_visits() const noexcept { return impl_->visits; }
HipTimings HipTreeSolver::timings() const { return impl_->device->timings(); }
void HipTreeSolver::reset_timings() { impl_->device->reset_timings(); }
void HipTreeSolver::set_timing(bool enabled) { impl_->device->set_timing(enabled); }

}  // namespace orrery::solvers
