// fg_emit.hip -- inflow emission of one fluid (IEmitter::update, whose
// reference implementation is a no-op, Emitter.h:46).  The particle array
// holds a reserve of inactive slots (w <= 0); every period of simulated time
// the next nu x nv layer of reserve slots is activated on a lattice in the
// inflow window, with the inflow velocity.  Velocities are in simulation
// units and a step moves a particle by v dt / simulation_scale world units
// (K_FluidSim.cu:433-438), so a layer has travelled `spacing` world units
// after period = spacing * simulation_scale / speed.  The step index lives
// on the device, so the emission is part of a captured step graph.
#include "fg_internal.hpp"

namespace fg {

// Number of layers emitted before step k: floor(k dt / period), in FP64.
__device__ __forceinline__ uint64_t layers_before(uint64_t k, double dt, double period) {
    return (uint64_t)floor((double)k * dt / period);
}

#pragma clang fp contract(off)
__global__ void __launch_bounds__(256) k_emit_inflow(float4* __restrict__ pos, float4* __restrict__ vel,
                                                     const uint32_t* __restrict__ step, fg_inflow_desc d, float dt,
                                                     double period) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t per = d.nu * d.nv;
    if (t >= per) return;
    const uint64_t k = step[0];
    const uint64_t l0 = layers_before(k, dt, period), l1 = layers_before(k + 1, dt, period);
    if (l1 == l0) return;  // no layer this step
    const uint64_t slot = (uint64_t)d.firstSlot + l0 * per + t;
    if (l0 >= d.numLayers || slot >= (uint64_t)d.firstSlot + (uint64_t)d.numLayers * per) return;
    const uint32_t iu = t % d.nu, iv = t / d.nu;
    const float fu = (float)iu * d.spacing, fv = (float)iv * d.spacing;
    pos[slot] = make_float4(d.origin.x + fu * d.axisU.x + fv * d.axisV.x, d.origin.y + fu * d.axisU.y + fv * d.axisV.y,
                            d.origin.z + fu * d.axisU.z + fv * d.axisV.z, 1.0f);
    vel[slot] = make_float4(d.speed * d.normal.x, d.speed * d.normal.y, d.speed * d.normal.z, 0.0f);
}
#pragma clang fp contract(fast)

__global__ void k_step_counter(uint32_t* __restrict__ step) { step[0] += 1u; }

}  // namespace fg

using namespace fg;

extern "C" {

double fg_inflow_period(const fg_inflow_desc* desc, float simulationScale) {
    if (!desc || !(desc->speed > 0.0f)) return 0.0;
    return (double)desc->spacing * (double)simulationScale / (double)desc->speed;
}

int fg_emit_inflow(fg_float4* positions, fg_float4* velocities, unsigned int* stepCounter, const fg_inflow_desc* desc,
                   float dt) {
    if (!positions || !velocities || !stepCounter || !desc) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_emit_inflow: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    const fg_inflow_desc& d = *desc;
    const uint64_t per = (uint64_t)d.nu * d.nv;
    if (per == 0 || per > 0xFFFFFFFFull || !(d.spacing > 0.0f) || !(d.speed > 0.0f) || !(dt > 0.0f) ||
        (uint64_t)d.firstSlot + per * d.numLayers > 0xFFFFFFFFull) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_emit_inflow: empty window, non-positive spacing/speed/dt or reserve "
                                           "beyond 2^32 slots");
        return FG_ERR_INVALID_ARGUMENT;
    }
    const State& st = state();
    if (!st.have_common) {
        set_error(FG_ERR_NOT_INITIALIZED, "fg_emit_inflow: call gInitializeCommonFluidParams first");
        return FG_ERR_NOT_INITIALIZED;
    }
    const double period = fg_inflow_period(desc, st.common.simulation_scale);
    // One layer at most per step: with dt > period a step would owe several
    // layers, all at the window, so such a step is refused.
    if (!(period > 0.0) || (double)dt > period) {
        set_error(FG_ERR_INVALID_ARGUMENT,
                  "fg_emit_inflow: dt %g exceeds the layer period %g (spacing * simulation_scale / speed); "
                  "lower dt or the speed, or widen the spacing",
                  (double)dt, period);
        return FG_ERR_INVALID_ARGUMENT;
    }
    hipLaunchKernelGGL(k_emit_inflow, dim3(grid_for(per, 256)), dim3(256), 0, st.stream, FG_F4(positions),
                       FG_F4(velocities), stepCounter, d, dt, period);
    hipLaunchKernelGGL(k_step_counter, dim3(1), dim3(1), 0, st.stream, stepCounter);
    return check_launch("fg_emit_inflow") ? FG_OK : fg_peek_error();
}

}  // extern "C"
