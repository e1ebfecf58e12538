// fg_batch.hip -- many independent fluids advanced with one launch per stage.
//
// Particles are scene-major and every scene owns a LINEAR grid whose buckets
// sit at [keyBase, keyBase + buckets) of one shared key space, so the single
// stable sort and the single cell-range pass of the one-fluid path serve all
// scenes.  Each lane handles one particle; a wavefront whose particles belong
// to several scenes runs once per distinct scene (readfirstlane loop), so the
// scene's parameters are wave-uniform and read through the scalar cache.
#include "fg_batch_bodies.hpp"
#include "fg_internal.hpp"
#include "fg_sph_fast.hpp"

namespace fg {

// Calls f(scene) with a wave-uniform scene index for every lane whose scene
// is below nscenes.
template <class F>
__device__ __forceinline__ void for_each_scene(uint32_t sc, uint32_t nscenes, F&& f) {
    if (sc >= nscenes) return;
    for (;;) {
        const uint32_t s0 = __builtin_amdgcn_readfirstlane(sc);
        if (sc == s0) {
            f(s0);
            return;
        }
    }
}

__device__ __forceinline__ GridDesc scene_grid(const fg_common_fluid_params& cp) {
    GridDesc g;
    g.vmin = make_float3(cp.volume_min.x, cp.volume_min.y, cp.volume_min.z);
    g.cs = make_float3(cp.cell_size.x, cp.cell_size.y, cp.cell_size.z);
    g.dims = make_int3(cp.bucket_size.x, cp.bucket_size.y, cp.bucket_size.z);
    g.nb = dev_num_buckets(cp);
    g.mode = FG_KEY_LINEAR;
    g.key_base = 0;
    return g;
}

__global__ void __launch_bounds__(256) k_batch_hash(const fg_scene_desc* __restrict__ scenes, uint32_t nscenes,
                                                    const uint32_t* __restrict__ scene_of,
                                                    const float4* __restrict__ pos, uint32_t* __restrict__ key,
                                                    uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t sc = scene_of[i];
    if (sc >= nscenes) {
        key[i] = kEmpty;  // sorted last; skipped by the cell-range pass
        return;
    }
    const float4 p = pos[i];
    for_each_scene(sc, nscenes, [&](uint32_t s) {
        const GridDesc g = scene_grid(scenes[s].common);
        const CellXYZ c = grid_cell(p.x, p.y, p.z, g);
        key[i] = scenes[s].keyBase + grid_linear_key(c.x, c.y, c.z, g);
    });
}

// The 27-cell stencil holds every neighbour only when each cell edge is at
// least the kernel radius.  The scene descriptors live in device memory, so
// the launchers cannot check them on the host (BatchedSimulator does, when
// it builds them): a scene with smaller cells gets NaN densities and
// velocities instead of sums that silently miss neighbours.
__device__ __forceinline__ bool scene_stencil_covers_h(const fg_common_fluid_params& cp) {
    const float s = cp.simulation_scale;
    return linear_reach(cp.kernel_radius, s, cp.cell_size.x) <= 1 &&
           linear_reach(cp.kernel_radius, s, cp.cell_size.y) <= 1 &&
           linear_reach(cp.kernel_radius, s, cp.cell_size.z) <= 1;
}

__device__ __forceinline__ void refuse_scene_particle(uint32_t i, const uint32_t* __restrict__ id, uint32_t n,
                                                      float4* __restrict__ vout) {
    const uint32_t o = id[i];
    if (o < n) vout[o] = make_float4(__builtin_nanf(""), __builtin_nanf(""), __builtin_nanf(""), 0.0f);
}

template <int BS, bool FLAT>
__global__ void __launch_bounds__(BS)
    k_batch_density(const fg_scene_desc* __restrict__ scenes, uint32_t nscenes, const uint32_t* __restrict__ scene_of,
                    const float4* __restrict__ sp, const float4* __restrict__ sv, float2* __restrict__ rho_p,
                    const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                    unsigned long long* __restrict__ masks, uint32_t n) {
    __shared__ uint32_t s_rb[FLAT ? 9 * BS : 1];
    __shared__ uint32_t s_re[FLAT ? 9 * BS : 1];
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if (i >= n) return;
    FlatRanges fr{s_rb, s_re, (uint32_t)BS, (uint32_t)threadIdx.x};
    for_each_scene(scene_of[i], nscenes, [&](uint32_t s) {
        const fg_scene_desc& d = scenes[s];
        if (!scene_stencil_covers_h(d.common)) {
            rho_p[i] = make_float2(__builtin_nanf(""), __builtin_nanf(""));
            return;
        }
        density_particle<true>(i, sp, sv, rho_p, start, end, n, d.gasConstant, d.restDensity, d.mass, d.common,
                               d.kernels, d.keyBase, masks, FLAT ? &fr : nullptr);
    });
}

template <int BS, bool FLAT>
__global__ void __launch_bounds__(BS)
    k_batch_forces(const fg_scene_desc* __restrict__ scenes, uint32_t nscenes, const uint32_t* __restrict__ scene_of,
                   const float4* __restrict__ sp, const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                   const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                   const unsigned long long* __restrict__ masks, float4* __restrict__ vout,
                   float4* __restrict__ pout, const uint32_t* __restrict__ id, uint32_t n) {
    __shared__ uint32_t s_rb[FLAT ? 9 * BS : 1];
    __shared__ uint32_t s_mm[FLAT ? 9 * BS : 1];
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if (i >= n) return;
    const Contacts none{};
    FlatRows fr{s_rb, s_mm, (uint32_t)BS, (uint32_t)threadIdx.x};
    for_each_scene(scene_of[i], nscenes, [&](uint32_t s) {
        const fg_scene_desc& d = scenes[s];
        if (!scene_stencil_covers_h(d.common)) return refuse_scene_particle(i, id, n, vout);
        forces_particle<true>(i, sp, sv, rho_p, start, end, vout, pout, id, n, d.mass, d.viscosity, d.common,
                              d.kernels, d.keyBase, masks, none, FLAT ? &fr : nullptr);
    });
}

// The same with each scene's bodies: scenes with bodies take the coupled
// force expression and contact against their body grid.
// amdgpu_waves_per_eu(5): 94-95 VGPRs without scratch (occupancy 4 without
// the hint; fitted to 6 waves it spills 17 VGPRs to scratch).
template <int BS, bool FLAT>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(5)))
    k_batch_forces_coupled(const fg_scene_desc* __restrict__ scenes, uint32_t nscenes,
                           const uint32_t* __restrict__ scene_of, const float4* __restrict__ sp,
                           const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                           const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                           const unsigned long long* __restrict__ masks, float4* __restrict__ vout,
                           float4* __restrict__ pout, const uint32_t* __restrict__ id, uint32_t n,
                           const BatchBodyScene* __restrict__ bodies) {
    __shared__ uint32_t s_rb[FLAT ? 9 * BS : 1];
    __shared__ uint32_t s_mm[FLAT ? 9 * BS : 1];
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if (i >= n) return;
    FlatRows fr{s_rb, s_mm, (uint32_t)BS, (uint32_t)threadIdx.x};
    for_each_scene(scene_of[i], nscenes, [&](uint32_t s) {
        const fg_scene_desc& d = scenes[s];
        if (!scene_stencil_covers_h(d.common)) return refuse_scene_particle(i, id, n, vout);
        Contacts c{};
        c.dyn = bodies[s].dyn;
        c.coupled = c.dyn.enabled;
        forces_particle<true, true>(i, sp, sv, rho_p, start, end, vout, pout, id, n, d.mass, d.viscosity, d.common,
                                    d.kernels, d.keyBase, masks, c, FLAT ? &fr : nullptr);
    });
}

#define FG_BATCH_LAUNCH_BS(KERN, BS, n, ...) \
    hipLaunchKernelGGL(KERN, dim3(grid_for(n, BS)), dim3(BS), 0, state().stream, __VA_ARGS__)

// Launches KERN<BS, FLAT> with BS = the tuned SPH block size.
#define FG_BATCH_LAUNCH2(KERN, FLAT, n, ...)                                                                \
    do {                                                                                                    \
        switch (state().sph_block) {                                                                        \
            case 64: FG_BATCH_LAUNCH_BS((KERN<64, FLAT>), 64, n, __VA_ARGS__); break;                       \
            case 128: FG_BATCH_LAUNCH_BS((KERN<128, FLAT>), 128, n, __VA_ARGS__); break;                    \
            case 512: FG_BATCH_LAUNCH_BS((KERN<512, FLAT>), 512, n, __VA_ARGS__); break;                    \
            default: FG_BATCH_LAUNCH_BS((KERN<256, FLAT>), 256, n, __VA_ARGS__);                           \
        }                                                                                                   \
    } while (0)

int batch_forces_coupled(const fg_scene_desc* scenes, uint32_t nscenes, const uint32_t* scene_of, float4* sp,
                         float4* sv, float2* rho_p, uint32_t* start, uint32_t* end, const unsigned long long* masks,
                         float4* vel, float4* pos, uint32_t* id, uint32_t n, const BatchBodyScene* bodies) {
    if (n == 0) return FG_OK;
    if (!scenes || !scene_of || !sp || !sv || !rho_p || !start || !end || !masks || !vel || !id || !bodies ||
        nscenes == 0 || (void*)vel == (void*)sv) {
        set_error(FG_ERR_INVALID_ARGUMENT, "batch_forces_coupled: null or aliased pointer, or no scenes");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (state().flat_forces)
        FG_BATCH_LAUNCH2(k_batch_forces_coupled, true, n, scenes, nscenes, scene_of, sp, sv, rho_p, start, end, masks,
                         vel, pos, id, n, bodies);
    else
        FG_BATCH_LAUNCH2(k_batch_forces_coupled, false, n, scenes, nscenes, scene_of, sp, sv, rho_p, start, end,
                         masks, vel, pos, id, n, bodies);
    return check_launch("batch_forces_coupled") ? FG_OK : fg_peek_error();
}

}  // namespace fg

using namespace fg;

extern "C" {

int fg_batch_hash(const fg_scene_desc* scenes, unsigned int numScenes, const unsigned int* sceneOf,
                  const fg_float4* positions, unsigned int* hashKey, unsigned int numParticles) {
    if (numParticles == 0) return FG_OK;
    if (!scenes || !sceneOf || !positions || !hashKey || numScenes == 0) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_batch_hash: null pointer or no scenes");
        return FG_ERR_INVALID_ARGUMENT;
    }
    hipLaunchKernelGGL(k_batch_hash, dim3(grid_for(numParticles, 256)), dim3(256), 0, state().stream, scenes,
                       numScenes, sceneOf, reinterpret_cast<const float4*>(positions), hashKey, numParticles);
    return check_launch("fg_batch_hash") ? FG_OK : fg_peek_error();
}

int fg_batch_density_nbr(const fg_scene_desc* scenes, unsigned int numScenes, const unsigned int* sceneOf,
                         fg_float4* orderedPositions, const fg_float4* orderedVelocity, fg_float2* mass_pressure,
                         unsigned int* gridStart, unsigned int* gridEnd, unsigned long long* neighborMasks,
                         unsigned int numParticles) {
    if (numParticles == 0) return FG_OK;
    if (!scenes || !sceneOf || !orderedPositions || !orderedVelocity || !mass_pressure || !gridStart || !gridEnd || !neighborMasks ||
        numScenes == 0) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_batch_density_nbr: null pointer or no scenes");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (state().flat_density)
        FG_BATCH_LAUNCH2(k_batch_density, true, numParticles, scenes, numScenes, sceneOf, FG_F4(orderedPositions),
                         reinterpret_cast<const float4*>(orderedVelocity), FG_F2(mass_pressure), gridStart, gridEnd, neighborMasks, numParticles);
    else
        FG_BATCH_LAUNCH2(k_batch_density, false, numParticles, scenes, numScenes, sceneOf, FG_F4(orderedPositions),
                         reinterpret_cast<const float4*>(orderedVelocity), FG_F2(mass_pressure), gridStart, gridEnd, neighborMasks, numParticles);
    return check_launch("fg_batch_density_nbr") ? FG_OK : fg_peek_error();
}

int fg_batch_forces_integrate_nbr(const fg_scene_desc* scenes, unsigned int numScenes, const unsigned int* sceneOf,
                                  fg_float4* orderedPositions, fg_float4* orderedVelocity,
                                  fg_float2* density_pressure, unsigned int* gridStart, unsigned int* gridEnd,
                                  const unsigned long long* neighborMasks, fg_float4* vel, fg_float4* pos,
                                  unsigned int* particleId, unsigned int numParticles) {
    if (numParticles == 0) return FG_OK;
    if (!scenes || !sceneOf || !orderedPositions || !orderedVelocity || !density_pressure || !gridStart ||
        !gridEnd || !neighborMasks || !vel || !particleId || numScenes == 0) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_batch_forces_integrate_nbr: null pointer or no scenes");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if ((void*)vel == (void*)orderedVelocity) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_batch_forces_integrate_nbr: vel must not alias orderedVelocity");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (state().flat_forces)
        FG_BATCH_LAUNCH2(k_batch_forces, true, numParticles, scenes, numScenes, sceneOf, FG_F4(orderedPositions),
                         FG_F4(orderedVelocity), FG_F2(density_pressure), gridStart, gridEnd, neighborMasks,
                         FG_F4(vel), pos ? FG_F4(pos) : nullptr, particleId, numParticles);
    else
        FG_BATCH_LAUNCH2(k_batch_forces, false, numParticles, scenes, numScenes, sceneOf, FG_F4(orderedPositions),
                         FG_F4(orderedVelocity), FG_F2(density_pressure), gridStart, gridEnd, neighborMasks,
                         FG_F4(vel), pos ? FG_F4(pos) : nullptr, particleId, numParticles);
    return check_launch("fg_batch_forces_integrate_nbr") ? FG_OK : fg_peek_error();
}

}  // extern "C"
