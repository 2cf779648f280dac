#pragma once
// Minimal restatement of the reference's AccelerationField
// (integrators/acceleration_field.hpp:37-70): the seam every integrator calls.
#include <span>

#include "orrery/core/types.hpp"
#include "orrery/core/vec3_span.hpp"

namespace orrery::integrators {

class AccelerationField {
public:
    virtual ~AccelerationField() = default;
    virtual void evaluate(core::Vec3Span<const core::Real> positions,
                          std::span<const core::Real> masses,
                          core::Vec3Span<core::Real> accelerations) = 0;

protected:
    AccelerationField() = default;
    AccelerationField(const AccelerationField&) = default;
    AccelerationField(AccelerationField&&) = default;
    AccelerationField& operator=(const AccelerationField&) = default;
    AccelerationField& operator=(AccelerationField&&) = default;
};

}  // namespace orrery::integrators
