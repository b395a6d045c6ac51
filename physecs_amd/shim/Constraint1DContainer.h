// Constraint1DContainer.h of the drop-in Scene: the part of the reference's header that a
// Joint subclass writes its rows through (include/Physecs/Constraint1DContainer.h:10-93
// Constraint1DView, :179-220 Constraint1DLayout and Constraint1DWriter), over the rows the
// GPU step takes from its joint callbacks (pgx_joint_row). A user joint's getSolverDesc
// declares rows with createConstraints<flags, count>() and its MakeConstraintsFunc fills
// them with next<flags>().setLinear(...).setC(...), as against the reference. Put this
// folder before the reference's include folder on the include path. The reference's
// containers themselves (Constraint1DContainer, its SIMD lanes) have no counterpart here:
// the device holds the rows.
#pragma once

#include <cstdint>
#include <vector>

#include <glm/glm.hpp>

#include "Constraint1D.h"   // ConstraintFlags (include/Physecs/Constraint1D.h:11-16)
#include "physecs_gpu.h"

namespace physecs {

// the setters of one row (include/Physecs/Constraint1DContainer.h:10-93)
template <int flags>
class Constraint1DView {
    pgx_joint_row* row;

public:
    explicit Constraint1DView(pgx_joint_row* row) : row(row) {}

    Constraint1DView& setLinear(const glm::vec3& linear) {
        row->linear[0] = linear.x; row->linear[1] = linear.y; row->linear[2] = linear.z;
        return *this;
    }
    Constraint1DView& setAngular0(const glm::vec3& angular0) {
        row->angular0[0] = angular0.x; row->angular0[1] = angular0.y; row->angular0[2] = angular0.z;
        return *this;
    }
    Constraint1DView& setAngular1(const glm::vec3& angular1) {
        row->angular1[0] = angular1.x; row->angular1[1] = angular1.y; row->angular1[2] = angular1.z;
        return *this;
    }
    Constraint1DView& setTargetVelocity(float targetVelocity) { row->target_velocity = targetVelocity; return *this; }
    Constraint1DView& setC(float c) { row->c = c; return *this; }
    Constraint1DView& setMin(float min) { row->min = min; return *this; }
    Constraint1DView& setMax(float max) { row->max = max; return *this; }
    Constraint1DView& setFrequency(float frequency) { row->frequency = frequency; return *this; }
    Constraint1DView& setDampingRatio(float dampingRatio) { row->damping_ratio = dampingRatio; return *this; }
};

// getSolverDesc's declaration of the joint's rows for this step, in creation order
// (include/Physecs/Constraint1DContainer.h:179-198)
class Constraint1DLayout {
    std::vector<int32_t>& flagsOut;

public:
    explicit Constraint1DLayout(std::vector<int32_t>& flagsOut) : flagsOut(flagsOut) {}

    template <int flags = NONE, int count = 1>
    void createConstraints() {
        for (int i = 0; i < count; ++i) flagsOut.push_back(flags);
    }
};

// MakeConstraintsFunc's access to the rows, one after another in creation order
// (include/Physecs/Constraint1DContainer.h:200-220). The flags argument names the row's
// type as in the reference; a call beyond the declared rows writes a row nobody reads.
class Constraint1DWriter {
    pgx_joint_row* rows;
    int n;
    int index = 0;
    pgx_joint_row spare{};

public:
    Constraint1DWriter(pgx_joint_row* rows, int n) : rows(rows), n(n) {}

    template <int flags = NONE>
    Constraint1DView<flags> next() {
        pgx_joint_row* r = index < n ? rows + index : &spare;
        ++index;
        return Constraint1DView<flags>(r);
    }
};

}  // namespace physecs
