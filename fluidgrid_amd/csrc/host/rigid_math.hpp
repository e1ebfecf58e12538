// rigid_math.hpp -- host-side set-up arithmetic shared by the single-fluid
// RigidBodySimulator and the batched simulator, so that both derive a body's
// inverse inertia and sphere offsets with the same operations.
#pragma once

#include <cmath>
#include <stdexcept>

namespace fg_host {

// 3x3 inverse in FP64 (row-major); throws when singular.
inline void invert3(const double* a, double* inv) {
    const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (!(std::fabs(det) > 0.0)) throw std::runtime_error("RigidBodySimulator: singular inertia tensor");
    const double id = 1.0 / det;
    inv[0] = c00 * id;
    inv[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    inv[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    inv[3] = c01 * id;
    inv[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    inv[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    inv[6] = c02 * id;
    inv[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    inv[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

}  // namespace fg_host
