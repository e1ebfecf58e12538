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

// Calls f(i) for the sorted particles of this workgroup: one block of BS
// particles (plain launch, blocks remapped per XCD), or, with a work queue,
// the items the workgroup takes from the counters until every slice is
// exhausted (every workgroup reaches the end: each counter only grows).
template <int BS, bool PERSIST, class F>
__device__ __forceinline__ void for_each_particle(const WorkQueue& q, uint32_t n, F&& f) {
    if (!PERSIST) {
        const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
        if (i < n) f(i);
        return;
    }
    __shared__ uint32_t s_item;
    const uint32_t span = (uint32_t)BS * q.per_item;
    const uint32_t items = (n + span - 1) / span;
    const uint32_t per = items / kQueueSlots, rem = items % kQueueSlots;
    const uint32_t x0 = blockIdx.x % kQueueSlots;
    for (uint32_t k = 0; k < kQueueSlots; k++) {
        const uint32_t x = (x0 + k) % kQueueSlots;
        const uint32_t lo = x * per + min(x, rem), cnt = per + (x < rem ? 1u : 0u);
        uint32_t* ctr = q.counters + x * kQueueStride;
        for (;;) {
            if (threadIdx.x == 0) s_item = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __syncthreads();
            const uint32_t it = s_item;
            __syncthreads();
            if (it >= cnt) break;
            const uint32_t base = (lo + it) * span;
            for (uint32_t b = 0; b < q.per_item; b++) {
                const uint32_t i = base + b * (uint32_t)BS + threadIdx.x;
                if (i < n) f(i);
            }
        }
    }
}

template <int BS, bool FLAT, bool PERSIST>
__global__ void __launch_bounds__(BS)
    k_batch_density(const fg_scene_desc* __restrict__ scenes, uint32_t nscenes, const uint32_t* __restrict__ scene_of,
                    const float4* __restrict__ sp, const float4* __restrict__ sv, float2* __restrict__ rho_p,
                    const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                    unsigned long long* __restrict__ masks, uint32_t n, WorkQueue q) {
    __shared__ uint32_t s_rb[FLAT ? 9 * BS : 1];
    __shared__ uint32_t s_re[FLAT ? 9 * BS : 1];
    __shared__ float4 s_stage[FLAT ? 1 : (BS / 64) * kStagePos];
    float4* stage = FLAT ? nullptr : s_stage + (threadIdx.x / 64) * kStagePos;
    FlatRanges fr{s_rb, s_re, (uint32_t)BS, (uint32_t)threadIdx.x};
    for_each_particle<BS, PERSIST>(q, n, [&](uint32_t i) {
        for_each_scene(scene_of[i], nscenes, [&](uint32_t s) {
            const fg_scene_desc& d = scenes[s];
            if (!scene_stencil_covers_h(d.common)) {
                rho_p[i] = make_float2(__builtin_nanf(""), __builtin_nanf(""));
                return;
            }
            density_particle<true>(i, sp, sv, rho_p, start, end, n, d.gasConstant, d.restDensity, d.mass, d.common,
                                   d.kernels, d.keyBase, masks, FLAT ? &fr : nullptr, stage);
        });
    });
}

template <int BS, bool FLAT, bool PERSIST>
__global__ void __launch_bounds__(BS)
    k_batch_forces(const fg_scene_desc* __restrict__ scenes, uint32_t nscenes, const uint32_t* __restrict__ scene_of,
                   const float4* __restrict__ sp, const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                   const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                   const unsigned long long* __restrict__ masks, float4* __restrict__ vout,
                   float4* __restrict__ pout, const uint32_t* __restrict__ id, uint32_t n, WorkQueue q) {
    __shared__ uint32_t s_rb[FLAT ? 9 * BS : 1];
    __shared__ uint32_t s_mm[FLAT ? 9 * BS : 1];
    __shared__ float4 s_stage[FLAT ? 1 : (BS / 64) * 2 * kStageRecs];
    float4* stage = FLAT ? nullptr : s_stage + (threadIdx.x / 64) * 2 * kStageRecs;
    const Contacts none{};
    FlatRows fr{s_rb, s_mm, (uint32_t)BS, (uint32_t)threadIdx.x};
    for_each_particle<BS, PERSIST>(q, n, [&](uint32_t i) {
        for_each_scene(scene_of[i], nscenes, [&](uint32_t s) {
            const fg_scene_desc& d = scenes[s];
            if (!scene_stencil_covers_h(d.common)) return refuse_scene_particle(i, id, n, vout);
            forces_particle<true>(i, sp, sv, rho_p, start, end, vout, pout, id, n, d.mass, d.viscosity, d.common,
                                  d.kernels, d.keyBase, masks, none, FLAT ? &fr : nullptr, nullptr, stage);
        });
    });
}

// The same with each scene's bodies: scenes with bodies take the coupled
// force expression and contact against their body grid.
// amdgpu_waves_per_eu(5): 94-95 VGPRs without scratch (occupancy 4 without
// the hint; fitted to 6 waves it spills 17 VGPRs to scratch).
template <int BS, bool FLAT, bool PERSIST>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(5)))
    k_batch_forces_coupled(const fg_scene_desc* __restrict__ scenes, uint32_t nscenes,
                           const uint32_t* __restrict__ scene_of, const float4* __restrict__ sp,
                           const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                           const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                           const unsigned long long* __restrict__ masks, float4* __restrict__ vout,
                           float4* __restrict__ pout, const uint32_t* __restrict__ id, uint32_t n,
                           const BatchBodyScene* __restrict__ bodies, WorkQueue q) {
    __shared__ uint32_t s_rb[FLAT ? 9 * BS : 1];
    __shared__ uint32_t s_mm[FLAT ? 9 * BS : 1];
    __shared__ float4 s_stage[FLAT ? 1 : (BS / 64) * 2 * kStageRecs];
    float4* stage = FLAT ? nullptr : s_stage + (threadIdx.x / 64) * 2 * kStageRecs;
    FlatRows fr{s_rb, s_mm, (uint32_t)BS, (uint32_t)threadIdx.x};
    for_each_particle<BS, PERSIST>(q, n, [&](uint32_t i) {
        for_each_scene(scene_of[i], nscenes, [&](uint32_t s) {
            const fg_scene_desc& d = scenes[s];
            if (!scene_stencil_covers_h(d.common)) return refuse_scene_particle(i, id, n, vout);
            Contacts c{};
            c.dyn = bodies[s].dyn;
            c.coupled = c.dyn.enabled;
            forces_particle<true, true>(i, sp, sv, rho_p, start, end, vout, pout, id, n, d.mass, d.viscosity,
                                        d.common, d.kernels, d.keyBase, masks, c, FLAT ? &fr : nullptr, nullptr,
                                        stage);
        });
    });
}

// The coupled force pass with body contact deferred (forces_particle with
// DEFER): without the contact walk it needs fewer registers than
// k_batch_forces_coupled.  amdgpu_waves_per_eu(7): 72 VGPRs without
// scratch (73 and occupancy 6 without the hint; 8 waves spill to scratch).
template <int BS, bool FLAT, bool PERSIST>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(7)))
    k_batch_forces_deferred(const fg_scene_desc* __restrict__ scenes, uint32_t nscenes,
                            const uint32_t* __restrict__ scene_of, const float4* __restrict__ sp,
                            const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                            const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                            const unsigned long long* __restrict__ masks, float4* __restrict__ vout,
                            float4* __restrict__ pout, const uint32_t* __restrict__ id, uint32_t n,
                            const BatchBodyScene* __restrict__ bodies, DeferredContact dc, WorkQueue q) {
    __shared__ uint32_t s_rb[FLAT ? 9 * BS : 1];
    __shared__ uint32_t s_mm[FLAT ? 9 * BS : 1];
    __shared__ float4 s_stage[FLAT ? 1 : (BS / 64) * 2 * kStageRecs];
    float4* stage = FLAT ? nullptr : s_stage + (threadIdx.x / 64) * 2 * kStageRecs;
    FlatRows fr{s_rb, s_mm, (uint32_t)BS, (uint32_t)threadIdx.x};
    for_each_particle<BS, PERSIST>(q, n, [&](uint32_t i) {
        for_each_scene(scene_of[i], nscenes, [&](uint32_t s) {
            const fg_scene_desc& d = scenes[s];
            if (!scene_stencil_covers_h(d.common)) return refuse_scene_particle(i, id, n, vout);
            Contacts c{};
            c.dyn = bodies[s].dyn;
            c.coupled = c.dyn.enabled;
            forces_particle<true, true, true>(i, sp, sv, rho_p, start, end, vout, pout, id, n, d.mass, d.viscosity,
                                              d.common, d.kernels, d.keyBase, masks, c, FLAT ? &fr : nullptr, &dc,
                                              stage);
        });
    });
}

// Second pass of the deferred form: contact, walls, clamp and integration
// of the deferred particles (finish_accel<true> and write_back, as
// forces_particle ends), from the pair sum in dc.acc.  Each deferred
// particle takes a pair of lanes that walk its stencil cells together
// (contact_cells): a body grid's walk is a chain of dependent loads (bucket
// range, then each sphere), and one lane per particle would walk its 27
// cells one after another.  Of 1, 2, 4, 8 and 16 lanes per particle, 2
// measured fastest (256 scenes with 479 bodies, 455k deferred particles).
constexpr uint32_t kContactLanes = 2;
__global__ void __launch_bounds__(256)
    k_batch_contact_finish(const fg_scene_desc* __restrict__ scenes, uint32_t nscenes,
                           const uint32_t* __restrict__ scene_of, const float4* __restrict__ sp,
                           const float4* __restrict__ sv, float4* __restrict__ vout, float4* __restrict__ pout,
                           const uint32_t* __restrict__ id, uint32_t n, const BatchBodyScene* __restrict__ bodies,
                           DeferredContact dc) {
    const uint32_t cnt = min(*dc.count, n);
    const uint32_t sub = threadIdx.x % kContactLanes;
    const uint32_t groups = gridDim.x * (blockDim.x / kContactLanes);
    // k is the same on every lane of a group, so a group is active or idle as a whole
    for (uint32_t k = (blockIdx.x * blockDim.x + threadIdx.x) / kContactLanes; k < cnt; k += groups) {
        const uint32_t i = dc.list[k];
        for_each_scene(scene_of[i], nscenes, [&](uint32_t s) {
            const BodyContact& dyn = bodies[s].dyn;
            const float4 pos = sp[i], vel = sv[i];
            // fluid_external with the dynamic grid only: +0 + its contact sum
            const F3 ext = dyn.enabled ? add(f3(0.0f, 0.0f, 0.0f), contact_cells<kContactLanes>(pos, vel, dyn, sub))
                                       : f3(0.0f, 0.0f, 0.0f);
            if (sub != 0) return;
            const fg_scene_desc& d = scenes[s];
            const float4 ia = dc.acc[i];
            const F3 a = finish_accel_ext(f3(ia.x, ia.y, ia.z), d.mass, pos, vel, d.common, dyn.enabled != 0, ext);
            write_back(id[i], n, pos, vel, a, d.common, vout, pout);
        });
    }
}

// Workgroups of kernel k (BS lanes) the GPU holds at once, a multiple of the
// queue's slices: the grid of a persistent launch.
static unsigned persistent_grid(const void* k, int bs) {
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, bs, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                hipSuccess || cus < 1)
        cus = 1;
    unsigned g = (unsigned)per_cu * (unsigned)cus;
    g -= g % kQueueSlots;
    return g < kQueueSlots ? kQueueSlots : g;
}

template <class K, class... A>
static void launch_batch(K kernel, int bs, uint32_t n, const WorkQueue& q, A... args) {
    hipStream_t st = state().stream;
    if (q.counters) {
        launch_fill_u32(q.counters, 0u, (uint64_t)kQueueSlots * kQueueStride, st);
        hipLaunchKernelGGL(kernel, dim3(persistent_grid(reinterpret_cast<const void*>(kernel), bs)), dim3(bs), 0, st,
                           args..., q);
    } else {
        hipLaunchKernelGGL(kernel, dim3(grid_for(n, bs)), dim3(bs), 0, st, args..., q);
    }
}

// Launches KERN<BS, FLAT, PERSIST> with BS = the tuned SPH block size and
// PERSIST = whether a work queue is given.
#define FG_BATCH_LAUNCH_P(KERN, BS, FLAT, n, q, ...)                                                        \
    do {                                                                                                    \
        if (q.counters) launch_batch(KERN<BS, FLAT, true>, BS, n, q, __VA_ARGS__);                          \
        else launch_batch(KERN<BS, FLAT, false>, BS, n, q, __VA_ARGS__);                                    \
    } while (0)
#define FG_BATCH_LAUNCH2(KERN, FLAT, n, q, ...)                                                             \
    do {                                                                                                    \
        switch (state().sph_block) {                                                                        \
            case 64: FG_BATCH_LAUNCH_P(KERN, 64, FLAT, n, q, __VA_ARGS__); break;                           \
            case 128: FG_BATCH_LAUNCH_P(KERN, 128, FLAT, n, q, __VA_ARGS__); break;                         \
            case 512: FG_BATCH_LAUNCH_P(KERN, 512, FLAT, n, q, __VA_ARGS__); break;                         \
            default: FG_BATCH_LAUNCH_P(KERN, 256, FLAT, n, q, __VA_ARGS__);                                 \
        }                                                                                                   \
    } while (0)

static bool queue_ok(const WorkQueue& q, const char* who) {
    if (q.counters && (q.per_item < 1 || q.per_item > 64)) {
        set_error(FG_ERR_INVALID_ARGUMENT, "%s: work item of %u blocks outside 1..64", who, q.per_item);
        return false;
    }
    return true;
}

int batch_density(const fg_scene_desc* scenes, uint32_t nscenes, const uint32_t* scene_of, float4* sp,
                  const float4* sv, float2* rho_p, uint32_t* start, uint32_t* end, unsigned long long* masks,
                  uint32_t n, WorkQueue q) {
    if (n == 0) return FG_OK;
    if (!scenes || !scene_of || !sp || !sv || !rho_p || !start || !end || !masks || nscenes == 0) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_batch_density_nbr: null pointer or no scenes");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (!queue_ok(q, "fg_batch_density_nbr")) return FG_ERR_INVALID_ARGUMENT;
    if (state().flat_density)
        FG_BATCH_LAUNCH2(k_batch_density, true, n, q, scenes, nscenes, scene_of, (const float4*)sp, sv, rho_p,
                         (const uint32_t*)start, (const uint32_t*)end, masks, n);
    else
        FG_BATCH_LAUNCH2(k_batch_density, false, n, q, scenes, nscenes, scene_of, (const float4*)sp, sv, rho_p,
                         (const uint32_t*)start, (const uint32_t*)end, masks, n);
    return check_launch("fg_batch_density_nbr") ? FG_OK : fg_peek_error();
}

int batch_forces(const fg_scene_desc* scenes, uint32_t nscenes, const uint32_t* scene_of, float4* sp, float4* sv,
                 float2* rho_p, uint32_t* start, uint32_t* end, const unsigned long long* masks, float4* vel,
                 float4* pos, uint32_t* id, uint32_t n, WorkQueue q) {
    if (n == 0) return FG_OK;
    if (!scenes || !scene_of || !sp || !sv || !rho_p || !start || !end || !masks || !vel || !id || nscenes == 0) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_batch_forces_integrate_nbr: null pointer or no scenes");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if ((void*)vel == (void*)sv) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_batch_forces_integrate_nbr: vel must not alias orderedVelocity");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (!queue_ok(q, "fg_batch_forces_integrate_nbr")) return FG_ERR_INVALID_ARGUMENT;
    if (state().flat_forces)
        FG_BATCH_LAUNCH2(k_batch_forces, true, n, q, scenes, nscenes, scene_of, (const float4*)sp, (const float4*)sv,
                         (const float2*)rho_p, (const uint32_t*)start, (const uint32_t*)end, masks, vel, pos,
                         (const uint32_t*)id, n);
    else
        FG_BATCH_LAUNCH2(k_batch_forces, false, n, q, scenes, nscenes, scene_of, (const float4*)sp, (const float4*)sv,
                         (const float2*)rho_p, (const uint32_t*)start, (const uint32_t*)end, masks, vel, pos,
                         (const uint32_t*)id, n);
    return check_launch("fg_batch_forces_integrate_nbr") ? FG_OK : fg_peek_error();
}

int batch_forces_coupled(const fg_scene_desc* scenes, uint32_t nscenes, const uint32_t* scene_of, float4* sp,
                         float4* sv, float2* rho_p, uint32_t* start, uint32_t* end, const unsigned long long* masks,
                         float4* vel, float4* pos, uint32_t* id, uint32_t n, const BatchBodyScene* bodies,
                         WorkQueue q) {
    if (n == 0) return FG_OK;
    if (!scenes || !scene_of || !sp || !sv || !rho_p || !start || !end || !masks || !vel || !id || !bodies ||
        nscenes == 0 || (void*)vel == (void*)sv) {
        set_error(FG_ERR_INVALID_ARGUMENT, "batch_forces_coupled: null or aliased pointer, or no scenes");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (!queue_ok(q, "batch_forces_coupled")) return FG_ERR_INVALID_ARGUMENT;
    if (state().flat_forces)
        FG_BATCH_LAUNCH2(k_batch_forces_coupled, true, n, q, scenes, nscenes, scene_of, (const float4*)sp,
                         (const float4*)sv, (const float2*)rho_p, (const uint32_t*)start, (const uint32_t*)end, masks,
                         vel, pos, (const uint32_t*)id, n, bodies);
    else
        FG_BATCH_LAUNCH2(k_batch_forces_coupled, false, n, q, scenes, nscenes, scene_of, (const float4*)sp,
                         (const float4*)sv, (const float2*)rho_p, (const uint32_t*)start, (const uint32_t*)end, masks,
                         vel, pos, (const uint32_t*)id, n, bodies);
    return check_launch("batch_forces_coupled") ? FG_OK : fg_peek_error();
}

int batch_forces_deferred(const fg_scene_desc* scenes, uint32_t nscenes, const uint32_t* scene_of, float4* sp,
                          float4* sv, float2* rho_p, uint32_t* start, uint32_t* end,
                          const unsigned long long* masks, float4* vel, float4* pos, uint32_t* id, uint32_t n,
                          const BatchBodyScene* bodies, DeferredContact dc, WorkQueue q) {
    if (n == 0) return FG_OK;
    if (!scenes || !scene_of || !sp || !sv || !rho_p || !start || !end || !masks || !vel || !id || !bodies ||
        nscenes == 0 || (void*)vel == (void*)sv || !dc.acc || !dc.list || !dc.count) {
        set_error(FG_ERR_INVALID_ARGUMENT, "batch_forces_deferred: null or aliased pointer, or no scenes");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (!queue_ok(q, "batch_forces_deferred")) return FG_ERR_INVALID_ARGUMENT;
    hipStream_t st = state().stream;
    launch_fill_u32(dc.count, 0u, 1, st);
    if (state().flat_forces)
        FG_BATCH_LAUNCH2(k_batch_forces_deferred, true, n, q, scenes, nscenes, scene_of, (const float4*)sp,
                         (const float4*)sv, (const float2*)rho_p, (const uint32_t*)start, (const uint32_t*)end, masks,
                         vel, pos, (const uint32_t*)id, n, bodies, dc);
    else
        FG_BATCH_LAUNCH2(k_batch_forces_deferred, false, n, q, scenes, nscenes, scene_of, (const float4*)sp,
                         (const float4*)sv, (const float2*)rho_p, (const uint32_t*)start, (const uint32_t*)end, masks,
                         vel, pos, (const uint32_t*)id, n, bodies, dc);
    // up to one group per particle (the count is known only on the device)
    unsigned blocks = grid_for((uint64_t)n * kContactLanes, 256);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_batch_contact_finish, dim3(blocks), dim3(256), 0, st, scenes, nscenes, scene_of,
                       (const float4*)sp, (const float4*)sv, vel, pos, (const uint32_t*)id, n, bodies, dc);
    return check_launch("batch_forces_deferred") ? FG_OK : fg_peek_error();
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
    return batch_density(scenes, numScenes, sceneOf, FG_F4(orderedPositions),
                         reinterpret_cast<const float4*>(orderedVelocity), FG_F2(mass_pressure), gridStart, gridEnd,
                         neighborMasks, numParticles, WorkQueue{nullptr, 1});
}

int fg_batch_forces_integrate_nbr(const fg_scene_desc* scenes, unsigned int numScenes, const unsigned int* sceneOf,
                                  fg_float4* orderedPositions, fg_float4* orderedVelocity,
                                  fg_float2* density_pressure, unsigned int* gridStart, unsigned int* gridEnd,
                                  const unsigned long long* neighborMasks, fg_float4* vel, fg_float4* pos,
                                  unsigned int* particleId, unsigned int numParticles) {
    return batch_forces(scenes, numScenes, sceneOf, FG_F4(orderedPositions), FG_F4(orderedVelocity),
                        FG_F2(density_pressure), gridStart, gridEnd, neighborMasks, FG_F4(vel),
                        pos ? FG_F4(pos) : nullptr, particleId, numParticles, WorkQueue{nullptr, 1});
}

}  // extern "C"
