// fg_walls.hpp -- penalty walls of the fluid box (GPU_Kernels/K_Boundaries_Walls.cu
// and K_Boundaries_Common.cu), with the constants the force kernel passes
// (K_FluidSim.cu:289-293): boundary distance 0.003043174, stiffness 20000,
// damping 256; a zero-coefficient friction term against the fixed box
// (-11,-40,-11)..(11,200,11).
#pragma once
#include "fg_internal.hpp"

namespace fg {
#pragma clang fp contract(off)

constexpr float kWallDistance = 0.003043174f;
constexpr float kWallStiffness = 20000.0f;
constexpr float kWallDamping = 256.0f;
constexpr float kWallEpsilon = 0.00001f;

// K_Boundaries_Common.cu:12-23
__device__ __forceinline__ F3 repulsion(F3 vel, F3 n, float dist) {
    return mul(kWallStiffness * dist - kWallDamping * dot(n, vel), n);
}

// K_Boundaries_Common.cu:48-82 with kinetic = static limit = 0
__device__ __forceinline__ F3 friction0(F3 vel, F3 n) {
    const F3 v_n = mul(vel, dot(n, vel));
    const F3 v_t = sub(vel, v_n);
    const F3 neg = f3(-v_t.x, -v_t.y, -v_t.z);
    if ((v_t.x + v_t.y + v_t.z) / 3.0f > 0.0f) return neg;
    return mul(0.0f, neg);
}

// K_Boundaries_Walls.cu:8-65, faces y-, y+, z-, z+, x-, x+
__device__ __forceinline__ F3 walls_no_penetration(F3 pos, F3 vel, const fg_common_fluid_params& cp) {
    const float s = cp.simulation_scale;
    F3 f = f3(0.0f, 0.0f, 0.0f);
    float d;
    d = kWallDistance - ((pos.y - cp.volume_min.y) * s);
    if (d > kWallEpsilon) f = add(f, repulsion(vel, f3(0, 1, 0), d));
    d = kWallDistance - ((cp.volume_max.y - pos.y) * s);
    if (d > kWallEpsilon) f = add(f, repulsion(vel, f3(0, -1, 0), d));
    d = kWallDistance - ((pos.z - cp.volume_min.z) * s);
    if (d > kWallEpsilon) f = add(f, repulsion(vel, f3(0, 0, 1), d));
    d = kWallDistance - ((cp.volume_max.z - pos.z) * s);
    if (d > kWallEpsilon) {
        const F3 n = f3(0, 0, -1);
        const float adj = kWallStiffness * d - kWallDamping * dot(n, vel);
        f = add(f, mul(adj, n));
    }
    d = kWallDistance - ((pos.x - cp.volume_min.x) * s);
    if (d > kWallEpsilon) f = add(f, repulsion(vel, f3(1, 0, 0), d));
    d = kWallDistance - ((cp.volume_max.x - pos.x) * s);
    if (d > kWallEpsilon) f = add(f, repulsion(vel, f3(-1, 0, 0), d));
    return f;
}

// K_Boundaries_Walls.cu:68-125 against the box the force kernel passes.
__device__ __forceinline__ F3 walls_no_slip(F3 pos, F3 vel, const fg_common_fluid_params& cp) {
    const float s = cp.simulation_scale;
    const F3 gmin = f3(-11.0f, -40.0f, -11.0f), gmax = f3(11.0f, 200.0f, 11.0f);
    F3 f = f3(0.0f, 0.0f, 0.0f);
    float d;
    d = kWallDistance - ((pos.y - gmin.y) * s);
    if (d > kWallEpsilon) f = add(f, friction0(vel, f3(0, 1, 0)));
    d = kWallDistance - ((gmax.y - pos.y) * s);
    if (d > kWallEpsilon) f = add(f, friction0(vel, f3(0, -1, 0)));
    d = kWallDistance - ((pos.z - gmin.z) * s);
    if (d > kWallEpsilon) f = add(f, friction0(vel, f3(0, 0, 1)));
    d = kWallDistance - ((gmax.z - pos.z) * s);
    if (d > kWallEpsilon) f = add(f, friction0(vel, f3(0, 0, -1)));
    d = kWallDistance - ((pos.x - gmin.x) * s);
    if (d > kWallEpsilon) f = add(f, friction0(vel, f3(1, 0, 0)));
    d = kWallDistance - ((gmax.x - pos.x) * s);
    if (d > kWallEpsilon) f = add(f, friction0(vel, f3(-1, 0, 0)));
    return f;
}

#pragma clang fp contract(fast)
}  // namespace fg
