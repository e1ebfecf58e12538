// fg_diag.hip -- FP64 global sums of one fluid for conservation reports:
// mass, linear momentum, kinetic energy, gravitational potential energy and
// the internal energy of the equation of state.  The reference has no such
// check; these sums are the global conservation checks of the step.
//
// Two launches: a fixed number of workgroups each reduce a grid-strided
// share in a fixed tree, then one workgroup adds the partial sums in order,
// so the result does not depend on timing.
#include <algorithm>

#include "fg_internal.hpp"

namespace fg {

constexpr int kDiagThreads = 256;
constexpr int kDiagBlocks = 512;
constexpr int kDiagQ = 8;

// q: 0 active particles, 1 mass, 2-4 momentum, 5 kinetic energy,
// 6 gravitational potential energy, 7 internal energy of the EOS.
__global__ void __launch_bounds__(kDiagThreads)
    k_conservation_part(const float4* __restrict__ pos, const float4* __restrict__ vel,
                        const float4* __restrict__ spos, const float2* __restrict__ rho_p, uint32_t n, double m,
                        double gx, double gy, double gz, double scale, double rho0, double k,
                        double* __restrict__ part) {
    __shared__ double s[kDiagQ][kDiagThreads];
    double a[kDiagQ] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t stride = gridDim.x * kDiagThreads;
    for (uint32_t i = blockIdx.x * kDiagThreads + threadIdx.x; i < n; i += stride) {
        const float4 p = pos[i];
        if (p.w > 0.0f) {
            const float4 v = vel[i];
            const double vx = v.x, vy = v.y, vz = v.z;
            a[0] += 1.0;
            a[1] += m;
            a[2] += m * vx;
            a[3] += m * vy;
            a[4] += m * vz;
            a[5] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
            a[6] -= m * scale * (gx * (double)p.x + gy * (double)p.y + gz * (double)p.z);
        }
        // the densities are in sorted order; their particles' flags are in spos
        if (spos[i].w > 0.0f) {
            const double r = rho_p[i].x;
            a[7] += m * k * (log(r / rho0) + rho0 / r - 1.0);
        }
    }
    for (int q = 0; q < kDiagQ; q++) s[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int w = kDiagThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < kDiagQ; q++) s[q][threadIdx.x] += s[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < kDiagQ) part[(size_t)blockIdx.x * kDiagQ + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void __launch_bounds__(kDiagThreads)
    k_conservation_final(const double* __restrict__ part, uint32_t nparts, double* __restrict__ out) {
    __shared__ double s[kDiagQ][kDiagThreads];
    double a[kDiagQ] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < nparts; b += kDiagThreads)
        for (int q = 0; q < kDiagQ; q++) a[q] += part[(size_t)b * kDiagQ + q];
    for (int q = 0; q < kDiagQ; q++) s[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int w = kDiagThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < kDiagQ; q++) s[q][threadIdx.x] += s[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < kDiagQ) out[threadIdx.x] = s[threadIdx.x][0];
}

}  // namespace fg

using namespace fg;

extern "C" {

size_t fg_conservation_workspace_size(void) { return sizeof(double) * kDiagQ * kDiagBlocks; }

int fg_conservation_f64(const fg_float4* positions, const fg_float4* velocities, const fg_float4* orderedPositions,
                        const fg_float2* density_pressure, unsigned int numParticles, float mass, float restDensity,
                        float gasConstant, double* out8, void* workspace, size_t workspace_bytes) {
    if (!positions || !velocities || !orderedPositions || !density_pressure || !out8 || !workspace) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_conservation_f64: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (workspace_bytes < fg_conservation_workspace_size()) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_conservation_f64: workspace too small");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (!(restDensity > 0.0f)) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_conservation_f64: rest density must be positive");
        return FG_ERR_INVALID_ARGUMENT;
    }
    const State& st = state();
    const fg_common_fluid_params& cp = st.common;
    double* part = static_cast<double*>(workspace);
    const uint32_t blocks = numParticles ? std::min<uint32_t>(kDiagBlocks, grid_for(numParticles, kDiagThreads)) : 1;
    hipLaunchKernelGGL(k_conservation_part, dim3(blocks), dim3(kDiagThreads), 0, st.stream,
                       reinterpret_cast<const float4*>(positions), reinterpret_cast<const float4*>(velocities),
                       reinterpret_cast<const float4*>(orderedPositions),
                       reinterpret_cast<const float2*>(density_pressure), numParticles, (double)mass,
                       (double)cp.global_external_force.x, (double)cp.global_external_force.y,
                       (double)cp.global_external_force.z, (double)cp.simulation_scale, (double)restDensity,
                       (double)gasConstant, part);
    hipLaunchKernelGGL(k_conservation_final, dim3(1), dim3(kDiagThreads), 0, st.stream, part, blocks, out8);
    return check_launch("fg_conservation_f64") ? FG_OK : fg_peek_error();
}

}  // extern "C"
