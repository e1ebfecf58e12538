// Emitter.cpp -- initial particle placement.
#include <algorithm>
#include <cstring>

#include "fluidgrid/GPUSPHFluid.h"

namespace Simulation {
namespace Emitter {

// Emitter.cpp:16-42: y-layers outermost, then z, then x, with the loop
// coordinates accumulated in float; remaining slots are inactive (w = -1).
void BlockEmitter::reset(float4* positions, int count) {
    int used = 0;
    float posY = mMin.y;
    bool any = (mMin.x < mMax.x) && (mMin.z < mMax.z) && mSpacing > 0.0f;
    while (any && used < count) {
        for (float z = mMin.z; z < mMax.z; z += mSpacing)
            for (float x = mMin.x; x < mMax.x; x += mSpacing)
                if (used < count) positions[used++] = make_float4(x, posY, z, 1.0f);
        posY += mSpacing;
    }
    for (; used < count; used++) positions[used] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
}

ArrayEmitter::ArrayEmitter(const float4* positions, const float4* velocities, int count)
    : mPos(positions, positions + count) {
    if (velocities) mVel.assign(velocities, velocities + count);
}

void ArrayEmitter::reset(float4* positions, int count) {
    const int m = std::min<int>(count, (int)mPos.size());
    if (m > 0) std::memcpy(positions, mPos.data(), sizeof(float4) * m);
    for (int i = m; i < count; i++) positions[i] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
}

}  // namespace Emitter
}  // namespace Simulation
