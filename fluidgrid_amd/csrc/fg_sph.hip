// fg_sph.hip -- SPH density, pressure/viscosity forces with penalty walls,
// and symplectic-Euler integration (the work of GPU_Kernels/K_FluidSim.cu and
// K_Boundaries_Walls.cu).
//
// Two arithmetic forms share the launchers:
//  * reference form (FG_MATH_REFERENCE, and always for FG_KEY_REF_XOR): one
//    lane per sorted particle walks the 27 stencil buckets exactly as
//    K_FluidSim.cu:46-79/127-205 do, with per-bucket partial sums and the
//    same per-pair expressions, contraction off;
//  * fast form (FG_MATH_FAST with FG_KEY_LINEAR): the three x-neighbour cells
//    of a stencil row are consecutive keys, so their particles form one
//    contiguous range of the sorted arrays; a row (or its end cells) is
//    skipped when its nearest point is farther than h; the pair loop keeps
//    only what depends on j (r^2 test, (h^2-r^2)^3 for density; 1/r, h-r and
//    1/rho_j for forces) and applies m*kPoly6, m/rho_i, kSpiky, kVisc*mu once.
//    The neighbour set and summation order within a row are those of the
//    reference; only rounding differs.
#include "fg_internal.hpp"
#include "fg_sph_fast.hpp"
#include "fg_sph_tile.hpp"
#include "fg_sph_mixed.hpp"

#include <mutex>
#include <unordered_map>

namespace fg {

constexpr int kSphThreads = 256;

// ---------------------------------------------------------------------------
// Reference-arithmetic kernels
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)

// Bucket of stencil offset (x,y,z) for a particle at p; false when the LINEAR
// cell lies outside the grid.
template <int MODE>
__device__ __forceinline__ bool stencil_bucket(const float4& p, const CellXYZ& c, int x, int y, int z,
                                               const fg_common_fluid_params& cp, uint32_t nb, uint32_t& b) {
    if (MODE == FG_KEY_REF_XOR) {
        const float ox = cp.cell_size.x * (float)x, oy = cp.cell_size.y * (float)y, oz = cp.cell_size.z * (float)z;
        b = hash_ref_xor(p.x + ox, p.y + oy, p.z + oz, cp.cell_size.x, cp.cell_size.y, cp.cell_size.z, nb);
        return true;
    } else {
        const int cx = c.x + x, cy = c.y + y, cz = c.z + z;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= cp.bucket_size.x || cy >= cp.bucket_size.y || cz >= cp.bucket_size.z)
            return false;
        b = linear_key(cx, cy, cz, cp);
        return true;
    }
}

// K_FluidSim.cu:210-254 with ProcessMassDensityGridD (:46-79)
template <int MODE>
__global__ void __launch_bounds__(kSphThreads)
    k_density_ref(const float4* __restrict__ sp, float2* __restrict__ rho_p, const uint32_t* __restrict__ start,
                  const uint32_t* __restrict__ end, uint32_t n, float k, float rho0, float m,
                  fg_common_fluid_params cp, fg_kernels_cache kc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = sp[i];
    if (!(p.w > 0.0f)) return;
    const float s = cp.simulation_scale, h = cp.kernel_radius, h2 = cp.kernel_radius_squared;
    const float mkp = m * kc.kPoly6;
    const uint32_t nb = dev_num_buckets(cp);
    CellXYZ c{0, 0, 0};
    if (MODE == FG_KEY_LINEAR) c = linear_cell(p.x, p.y, p.z, cp);
    float rho = 0.0f;
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                uint32_t b;
                if (!stencil_bucket<MODE>(p, c, x, y, z, cp, nb, b)) continue;
                float part = 0.0f;
                const uint32_t s0 = start[b];
                if (s0 != kEmpty) {
                    const F3 p1 = mul(xyz(p), s);
                    const uint32_t e = min(end[b], n);
                    for (uint32_t j = s0; j < e; j++) {
                        const F3 p2 = mul(xyz(sp[j]), s);
                        const F3 d = sub(p1, p2);
                        const float l = len(d);
                        if (l < h) part = part + mkp * powf(h2 - (l * l), 3.0f);
                    }
                }
                rho = rho + part;
            }
    rho_p[i] = make_float2(fmaxf(1.0f, rho), k * (rho - rho0));
}

// Pressure + viscosity between particle i and the particles of one bucket
// (K_FluidSim.cu:85-172).
__device__ __forceinline__ F3 bucket_internal_ref(uint32_t i, uint32_t s0, uint32_t e, const float4& pos,
                                                  const float4& vel, float2 dp1, const float4* __restrict__ sp,
                                                  const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                                                  float s, float h, float kSpiky, float kVisc, float mu) {
    F3 total = f3(0.0f, 0.0f, 0.0f);
    const F3 p1 = mul(xyz(pos), s);
    const F3 v1 = xyz(vel);
    for (uint32_t j = s0; j < e; j++) {
        if (j == i) continue;
        const F3 p2 = mul(xyz(sp[j]), s);
        const float2 dp2 = rho_p[j];
        const F3 v2 = xyz(sv[j]);
        // SPHCalculatePressureForceD, K_FluidSim.cu:107-120
        F3 fp = f3(0.0f, 0.0f, 0.0f);
        {
            const float pp = (dp1.y + dp2.y) / (dp1.x * dp2.x);
            F3 diff = sub(p1, p2);
            const float l = len(diff);
            if (!(l > h)) {
                const float r = kSpiky * ((h - l) * (h - l));
                fp = mul(mul(mul(divs(diff, l), r), pp), -1.0f);
            }
        }
        total = add(total, fp);
        // SPHCalculateViscosityForceD, K_FluidSim.cu:85-102
        F3 fv = f3(0.0f, 0.0f, 0.0f);
        {
            const F3 diff = sub(p1, p2);
            const float l = len(diff);
            if (!(l > h)) {
                const float lap = kVisc * (h - l);
                const float rr = dp1.x * dp2.x;
                fv = f3(((v2.x - v1.x) / rr) * lap * mu, ((v2.y - v1.y) / rr) * lap * mu,
                        ((v2.z - v1.z) / rr) * lap * mu);
            }
        }
        total = add(total, fv);
    }
    return total;
}

// K_FluidSim.cu:260-309 (+ fused integrate when pout != nullptr); with
// COUPLED, K_FluidSim.cu:315-413.
template <int MODE, bool COUPLED>
__global__ void __launch_bounds__(kSphThreads)
    k_forces_ref(const float4* __restrict__ sp, const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                 const uint32_t* __restrict__ start, const uint32_t* __restrict__ end, float4* __restrict__ vout,
                 float4* __restrict__ pout, const uint32_t* __restrict__ id, uint32_t n, float m, float mu,
                 fg_common_fluid_params cp, fg_kernels_cache kc, Contacts bc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 pos = sp[i];
    const float4 vel = sv[i];
    F3 a = f3(0.0f, 0.0f, 0.0f);
    if (pos.w > 0.0f) {
        const float2 dp1 = rho_p[i];
        const uint32_t nb = dev_num_buckets(cp);
        CellXYZ c{0, 0, 0};
        if (MODE == FG_KEY_LINEAR) c = linear_cell(pos.x, pos.y, pos.z, cp);
        F3 internal = f3(0.0f, 0.0f, 0.0f);
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++) {
                    uint32_t b;
                    if (!stencil_bucket<MODE>(pos, c, x, y, z, cp, nb, b)) continue;
                    const uint32_t s0 = start[b];
                    if (s0 == kEmpty) continue;
                    internal = add(internal, bucket_internal_ref(i, s0, min(end[b], n), pos, vel, dp1, sp, sv, rho_p,
                                                                 cp.simulation_scale, cp.kernel_radius,
                                                                 kc.kSpikyGradient, kc.kViscosityLaplacian, mu));
                }
        a = finish_accel<COUPLED>(internal, m, pos, vel, cp, bc);
    }
    write_back(id[i], n, pos, vel, a, cp, vout, pout);
}

// K_FluidSim.cu:418-438
__global__ void __launch_bounds__(256) k_integrate(float4* __restrict__ pos, const float4* __restrict__ vel,
                                                   uint32_t n, float q) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    if (p.w > 0.0f) {
        const float4 v = vel[i];
        pos[i] = make_float4(p.x + v.x * q, p.y + v.y * q, p.z + v.z * q, p.w + v.w * q);
    }
}

#pragma clang fp contract(fast)

// ---------------------------------------------------------------------------
// Fast LINEAR kernels (per-particle work in fg_sph_fast.hpp)
// ---------------------------------------------------------------------------

template <bool MASKS, int BS>
__global__ void __launch_bounds__(BS)
    k_density_fast(const float4* __restrict__ sp, const float4* __restrict__ sv, float2* __restrict__ rho_p,
                   const uint32_t* __restrict__ start,
                   const uint32_t* __restrict__ end, uint32_t n, float k, float rho0, float m,
                   fg_common_fluid_params cp, fg_kernels_cache kc, unsigned long long* __restrict__ masks) {
    // one LDS slice of kStagePos positions per wavefront (staged row loop)
    __shared__ float4 s_stage[(BS / 64) * kStagePos];
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if (i >= n) return;
    density_particle<MASKS>(i, sp, sv, rho_p, start, end, n, k, rho0, m, cp, kc, 0u, masks, nullptr,
                            s_stage + (threadIdx.x / 64) * kStagePos);
}

// Flat schedule (row table in LDS); same results as k_density_fast<MASKS>.
template <bool MASKS, int BS>
__global__ void __launch_bounds__(BS)
    k_density_flat(const float4* __restrict__ sp, const float4* __restrict__ sv, float2* __restrict__ rho_p,
                   const uint32_t* __restrict__ start,
                   const uint32_t* __restrict__ end, uint32_t n, float k, float rho0, float m,
                   fg_common_fluid_params cp, fg_kernels_cache kc, unsigned long long* __restrict__ masks) {
    __shared__ uint32_t s_rb[9 * BS];
    __shared__ uint32_t s_re[9 * BS];
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if (i >= n) return;
    FlatRanges fr{s_rb, s_re, (uint32_t)BS, (uint32_t)threadIdx.x};
    density_particle<MASKS>(i, sp, sv, rho_p, start, end, n, k, rho0, m, cp, kc, 0u, masks, &fr);
}

// amdgpu_waves_per_eu(7): the masked kernels load each stencil row's cell
// bounds and mask a row ahead (forces_rows_staged<true>), which needs about
// 70 registers: at 8 waves per SIMD that spills to scratch and the pass is
// 25% slower, at 7 it is 4-7% faster than the row-by-row loads at 8.  The
// hint also holds the coupled variant, whose two body-grid descriptors take
// scalar registers (106 SGPRs without it), to 7 waves.
template <bool MASKS, bool COUPLED, int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(7)))
    k_forces_fast(const float4* __restrict__ sp, const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                  const uint32_t* __restrict__ start, const uint32_t* __restrict__ end, float4* __restrict__ vout,
                  float4* __restrict__ pout, const uint32_t* __restrict__ id, uint32_t n, float m, float mu,
                  fg_common_fluid_params cp, fg_kernels_cache kc, const unsigned long long* __restrict__ masks,
                  Contacts bc) {
    // one LDS slice of kStageRecs neighbour records per wavefront (staged row loop)
    __shared__ float4 s_stage[MASKS ? (BS / 64) * 2 * kStageRecs : 1];
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if (i >= n) return;
    forces_particle<MASKS, COUPLED, false, true>(i, sp, sv, rho_p, start, end, vout, pout, id, n, m, mu, cp, kc, 0u,
                                                 masks, bc, nullptr, nullptr,
                                                 MASKS ? s_stage + (threadIdx.x / 64) * 2 * kStageRecs : nullptr);
}

// Flat schedule (row table in LDS); same results as k_forces_fast<true, COUPLED>.
// amdgpu_waves_per_eu(8): the coupled variant's two body-grid descriptors
// take scalar registers (106 SGPRs without the hint, occupancy 7).
template <bool COUPLED, int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(8)))
    k_forces_flat(const float4* __restrict__ sp, const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                  const uint32_t* __restrict__ start, const uint32_t* __restrict__ end, float4* __restrict__ vout,
                  float4* __restrict__ pout, const uint32_t* __restrict__ id, uint32_t n, float m, float mu,
                  fg_common_fluid_params cp, fg_kernels_cache kc, const unsigned long long* __restrict__ masks,
                  Contacts bc) {
    __shared__ uint32_t s_rb[9 * BS];
    __shared__ uint32_t s_mm[9 * BS];
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if (i >= n) return;
    FlatRows fr{s_rb, s_mm, (uint32_t)BS, (uint32_t)threadIdx.x};
    forces_particle<true, COUPLED>(i, sp, sv, rho_p, start, end, vout, pout, id, n, m, mu, cp, kc, 0u, masks, bc,
                                   &fr);
}

// The coupled masked force pass with body contact deferred: a particle
// inside a body grid's expanded bounds stores its pair sum and joins a list
// (forces_particle with DEFER); the others finish here without the contact
// walk, which leaves the wavefronts that never meet a body free of it.  Row
// loop with row-ahead loads, as k_forces_fast<true, true>.
template <int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(7)))
    k_forces_deferred(const float4* __restrict__ sp, const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                      const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                      float4* __restrict__ vout, float4* __restrict__ pout, const uint32_t* __restrict__ id,
                      uint32_t n, float m, float mu, fg_common_fluid_params cp, fg_kernels_cache kc,
                      const unsigned long long* __restrict__ masks, Contacts bc, DeferredContact dc) {
    __shared__ float4 s_stage[(BS / 64) * 2 * kStageRecs];
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if (i >= n) return;
    forces_particle<true, true, true, true>(i, sp, sv, rho_p, start, end, vout, pout, id, n, m, mu, cp, kc, 0u, masks,
                                            bc, nullptr, &dc, s_stage + (threadIdx.x / 64) * 2 * kStageRecs);
}

// Second pass of the deferred form: body contact (fluid_external: the
// static grid, then the dynamic one, summed from zero), walls, clamp and
// integration of the listed particles from their pair sums.  Each particle
// takes a pair of lanes that walk its body-grid cells together
// (contact_cells, bit-identical to the one-lane walk).
constexpr uint32_t kFinishLanes = 2;
__global__ void __launch_bounds__(256)
    k_contact_finish(const float4* __restrict__ sp, const float4* __restrict__ sv, float4* __restrict__ vout,
                     float4* __restrict__ pout, const uint32_t* __restrict__ id, uint32_t n, float m,
                     fg_common_fluid_params cp, Contacts bc, DeferredContact dc) {
    const uint32_t cnt = min(*dc.count, n);
    const uint32_t sub = threadIdx.x % kFinishLanes;
    const uint32_t groups = gridDim.x * (blockDim.x / kFinishLanes);
    // k is the same on both lanes of a group, so a group is active or idle as a whole
    for (uint32_t k = (blockIdx.x * blockDim.x + threadIdx.x) / kFinishLanes; k < cnt; k += groups) {
        const uint32_t i = dc.list[k];
        const float4 pos = sp[i], vel = sv[i];
        F3 e = f3(0.0f, 0.0f, 0.0f);
        if (bc.st.enabled) e = add(e, contact_cells<kFinishLanes>(pos, vel, bc.st, sub));
        if (bc.dyn.enabled) e = add(e, contact_cells<kFinishLanes>(pos, vel, bc.dyn, sub));
        if (sub != 0) continue;
        const float4 ia = dc.acc[i];
        const F3 a = finish_accel_ext(f3(ia.x, ia.y, ia.z), m, pos, vel, cp, bc.coupled != 0, e);
        write_back(id[i], n, pos, vel, a, cp, vout, pout);
    }
}

// Mixed-precision density (fg_sph_mixed.hpp): the FP32 pass's masks and
// records, FP16-culled candidates, compensated sums.
template <int BS>
__global__ void __launch_bounds__(BS)
    k_density_mixed(const float4* __restrict__ sp, const float4* __restrict__ sv, float2* __restrict__ rho_p,
                    const uint32_t* __restrict__ start, const uint32_t* __restrict__ end, uint32_t n, float k,
                    float rho0, float m, fg_common_fluid_params cp, fg_kernels_cache kc,
                    unsigned long long* __restrict__ masks) {
    // per wavefront: kMixedStage + 4 pair records, then kMixedStage positions
    __shared__ float4 s_stage[(BS / 64) * (2 * kMixedStage + 4)];
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x;
    if (i >= n) return;
    density_particle_mixed(i, sp, sv, rho_p, start, end, n, k, rho0, m, cp, kc, masks,
                           s_stage + (threadIdx.x / 64) * (2 * kMixedStage + 4));
}

// LDS-tiled passes (fg_sph_tile.hpp).  Tile b of the grid is the sorted
// particles [b * kTileLanes, (b + 1) * kTileLanes); blocks are remapped so
// each XCD walks a contiguous range of tiles.
__global__ void __launch_bounds__(kTileLanes)
    k_density_tile(const float4* __restrict__ sp, const float4* __restrict__ sv, float2* __restrict__ rho_p,
                   const uint32_t* __restrict__ start, const uint32_t* __restrict__ end, uint32_t n, float k,
                   float rho0, float m, fg_common_fluid_params cp, fg_kernels_cache kc,
                   unsigned long long* __restrict__ masks) {
    __shared__ TileGeom g;
    __shared__ TileSetupLds L;
    __shared__ float4 lpos[kTileCap];
    __shared__ __align__(16) uint16_t llists[kTileLanes * kListMax];
    const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = tile * kTileLanes + threadIdx.x;
    const bool have = i < n;
    const uint32_t nvalid = min(n - tile * kTileLanes, (uint32_t)kTileLanes);
    const float4 p = have ? sp[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 v = have ? sv[i] : p;  // (loaded before the set-up, whose latency it shares)
    const uint32_t fb = tile_setup(g, L, n, nvalid, have, p, cp, 0u, start, end);
    // the staging plan, for the force pass
    if (threadIdx.x < kTileDescWords) {
        const uint32_t t = threadIdx.x, nseg = fb ? 0u : 9u * g.nrows;
        uint32_t v = 0u;
        if (t == 0) v = fb;
        else if (t == 1) v = fb ? 0u : g.nrows;
        else if (t == 2) v = fb ? 0u : g.total;
        else if (t >= 4 && t < 4 + (uint32_t)kTileSegs) v = t - 4 < nseg ? g.sb[t - 4] : 0u;
        else if (t >= 4 + (uint32_t)kTileSegs && t < 4 + 2 * (uint32_t)kTileSegs)
            v = t - 4 - kTileSegs < nseg ? g.len[t - 4 - kTileSegs] : 0u;
        else if (t >= 4 + 2 * (uint32_t)kTileSegs && t < 4 + 3 * (uint32_t)kTileSegs)
            v = t - 4 - 2 * kTileSegs < nseg ? g.off[t - 4 - 2 * kTileSegs] : 0u;
        nbr_descs(masks, n)[(size_t)tile * kTileDescWords + t] = v;
    }
    if (fb) {
        if (have) {
            density_particle<true>(i, sp, sv, rho_p, start, end, n, k, rho0, m, cp, kc, 0u, masks);
            nbr_counts(masks, n)[i] = kListOverflow;
        }
        return;
    }
    tile_fill(g, sp, 1, lpos);
    tile_fill_wait();
    if (have)
        tile_density_particle(g, lpos, llists + threadIdx.x * kListMax, i, p, v, rho_p, start, end, n, k, rho0, m,
                              cp, kc, masks);
}

template <bool COUPLED>
__global__ void __launch_bounds__(kTileLanes)
    k_forces_tile(const float4* __restrict__ sp, const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                  const uint32_t* __restrict__ start, const uint32_t* __restrict__ end, float4* __restrict__ vout,
                  float4* __restrict__ pout, const uint32_t* __restrict__ id, uint32_t n, float m, float mu,
                  fg_common_fluid_params cp, fg_kernels_cache kc, const unsigned long long* __restrict__ masks,
                  Contacts bc) {
    __shared__ TileGeom g;
    __shared__ __align__(16) float4 recs[2 * kTileCap];
    float4* recA = recs;
    float4* recB = recs + kTileCap;
    const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = tile * kTileLanes + threadIdx.x;
    const bool have = i < n;
    // Everything a lane needs besides the staged records is loaded first, so
    // its latency overlaps the descriptor load and the staging.
    float4 pos = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vel = pos;
    float2 dpi = make_float2(1.0f, 0.0f);
    uint32_t cnt = 0u, oid = n;
    uint4 l0 = make_uint4(0u, 0u, 0u, 0u), l1 = l0;
    if (have) {
        const uint4* lists = nbr_lists(masks, n);
        pos = sp[i];
        vel = sv[i];
        dpi = rho_p[i];
        cnt = nbr_counts(masks, n)[i];
        oid = id[i];
        l0 = lists[i];
        l1 = lists[(size_t)n + i];
    }
    // the staging plan the density pass wrote for this tile
    if (threadIdx.x < kTileDescWords) {
        const uint32_t t = threadIdx.x;
        const uint32_t v = nbr_descs(masks, n)[(size_t)tile * kTileDescWords + t];
        if (t == 0) g.fallback = v;
        else if (t == 1) g.nrows = v;
        else if (t == 2) g.total = v;
        else if (t >= 4 && t < 4 + (uint32_t)kTileSegs) g.sb[t - 4] = v;
        else if (t >= 4 + (uint32_t)kTileSegs && t < 4 + 2 * (uint32_t)kTileSegs) g.len[t - 4 - kTileSegs] = v;
        else if (t >= 4 + 2 * (uint32_t)kTileSegs && t < 4 + 3 * (uint32_t)kTileSegs)
            g.off[t - 4 - 2 * kTileSegs] = v;
    }
    __syncthreads();
    if (g.fallback != 0u || g.total > kTileCap || g.nrows > (uint32_t)kTileRowsMax) {
        if (have)
            forces_particle<true, COUPLED>(i, sp, sv, rho_p, start, end, vout, pout, id, n, m, mu, cp, kc, 0u, masks,
                                           bc);
        return;
    }
    const float4* rec = nbr_records(masks, n);
    tile_fill(g, rec, 2, recA);
    tile_fill(g, rec + 1, 2, recB);
    tile_fill_wait();
    if (!have) return;
    F3 a = f3(0.0f, 0.0f, 0.0f);
    if (pos.w > 0.0f) {
        if (cnt > kListMax) cnt = kListOverflow;  // (a count the density pass cannot write reads the masks)
        a = tile_forces_particle<COUPLED>(recA, recB, i, pos, vel, dpi, cnt, l0, l1, sp, sv, rho_p, start, end, n, m,
                                          mu, cp, kc, masks, bc);
    }
    write_back(oid, n, pos, vel, a, cp, vout, pout);
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------

// Which form the last density pass wrote into a neighbour buffer: lists (the
// tiled pass) or row masks only.  The force pass follows it, whatever the
// schedule is when it runs, so the two passes always agree on the format.
static std::mutex g_nbr_mu;
static std::unordered_map<const void*, bool>& nbr_tiled() {
    static std::unordered_map<const void*, bool> m;
    return m;
}
static void note_nbr_format(const void* masks, bool tiled) {
    std::lock_guard<std::mutex> lock(g_nbr_mu);
    nbr_tiled()[masks] = tiled;
}
static bool nbr_format_tiled(const void* masks) {
    std::lock_guard<std::mutex> lock(g_nbr_mu);
    auto it = nbr_tiled().find(masks);
    return it != nbr_tiled().end() && it->second;
}
// Neighbour buffers whose FP16 pair records a cell pass has written, with
// the particle count it wrote them for (the records' place depends on it).
static std::unordered_map<const void*, uint32_t>& half_written() {
    static std::unordered_map<const void*, uint32_t> m;
    return m;
}
void note_half_records(const void* buf, uint32_t n) {
    std::lock_guard<std::mutex> lock(g_nbr_mu);
    half_written()[buf] = n;
}
static bool half_records_for(const void* buf, uint32_t n) {
    std::lock_guard<std::mutex> lock(g_nbr_mu);
    auto it = half_written().find(buf);
    return it != half_written().end() && it->second == n;
}
static bool have_params(const char* who) {
    if (!state().have_common || !state().have_kernels) {
        set_error(FG_ERR_NOT_INITIALIZED, "%s: call gInitializeCommonFluidParams and gInitializeKernelParams first",
                  who);
        return false;
    }
    return true;
}

// The density and force passes walk the 27 cells around a particle's own,
// which holds every neighbour only when the cells are at least h on every
// axis; smaller cells are refused rather than summed short.
static bool stencil_covers_h(const char* who) {
    const fg_common_fluid_params& cp = state().common;
    const float s = cp.simulation_scale;
    if (linear_reach(cp.kernel_radius, s, cp.cell_size.x) > 1 ||
        linear_reach(cp.kernel_radius, s, cp.cell_size.y) > 1 ||
        linear_reach(cp.kernel_radius, s, cp.cell_size.z) > 1) {
        set_error(FG_ERR_UNSUPPORTED, "%s: cell_size is smaller than the kernel radius (the 27-cell stencil "
                  "would miss neighbours)", who);
        return false;
    }
    return true;
}

static bool use_fast() { return state().key_mode == FG_KEY_LINEAR && state().math_mode == FG_MATH_FAST; }

#define FG_SPH_LAUNCH_BS(KERN, BS, n, ...) \
    hipLaunchKernelGGL(KERN, dim3(grid_for(n, BS)), dim3(BS), 0, state().stream, __VA_ARGS__)

// Launch KERN<A, BS> / KERN<A, B, BS> with BS = the tuned SPH block size.
#define FG_SPH_LAUNCH(KERN, A, n, ...)                                                                      \
    do {                                                                                                    \
        switch (state().sph_block) {                                                                        \
            case 64: FG_SPH_LAUNCH_BS((KERN<A, 64>), 64, n, __VA_ARGS__); break;                            \
            case 128: FG_SPH_LAUNCH_BS((KERN<A, 128>), 128, n, __VA_ARGS__); break;                         \
            case 512: FG_SPH_LAUNCH_BS((KERN<A, 512>), 512, n, __VA_ARGS__); break;                         \
            default: FG_SPH_LAUNCH_BS((KERN<A, 256>), 256, n, __VA_ARGS__);                                \
        }                                                                                                   \
    } while (0)

#define FG_SPH_LAUNCH2(KERN, A, B, n, ...)                                                                  \
    do {                                                                                                    \
        switch (state().sph_block) {                                                                        \
            case 64: FG_SPH_LAUNCH_BS((KERN<A, B, 64>), 64, n, __VA_ARGS__); break;                         \
            case 128: FG_SPH_LAUNCH_BS((KERN<A, B, 128>), 128, n, __VA_ARGS__); break;                      \
            case 512: FG_SPH_LAUNCH_BS((KERN<A, B, 512>), 512, n, __VA_ARGS__); break;                      \
            default: FG_SPH_LAUNCH_BS((KERN<A, B, 256>), 256, n, __VA_ARGS__);                             \
        }                                                                                                   \
    } while (0)

int launch_density(float4* sp, float2* rho_p, uint32_t* start, uint32_t* end, uint32_t n, float k, float rho0,
                   float m, unsigned long long* masks = nullptr, const float4* sv = nullptr) {
    if (!stencil_covers_h("density pass")) return FG_ERR_UNSUPPORTED;
    const State& st = state();
    const dim3 grid(grid_for(n, kSphThreads)), block(kSphThreads);
    const bool mixed = st.precision == FG_PRECISION_MIXED && use_fast() && masks;
    if (mixed && !half_records_for(masks, n)) {
        set_error(FG_ERR_NOT_INITIALIZED, "density pass (mixed precision): no cell pass has written FP16 offsets "
                  "for %u particles into this neighbour buffer (fg_cell_offsets_pos_vel with neighborBuffer)", n);
        return FG_ERR_NOT_INITIALIZED;
    }
    if (masks) note_nbr_format(masks, use_fast() && st.tiled && !mixed);
    if (mixed)
        FG_SPH_LAUNCH_BS(k_density_mixed<256>, 256, n, sp, sv, rho_p, start, end, n, k, rho0, m, st.common,
                         st.kernels, masks);
    else if (use_fast() && masks && st.tiled)
        hipLaunchKernelGGL(k_density_tile, dim3(grid_for(n, kTileLanes)), dim3(kTileLanes), 0, st.stream, sp, sv,
                           rho_p, start, end, n, k, rho0, m, st.common, st.kernels, masks);
    else if (use_fast() && masks && st.flat_density)
        FG_SPH_LAUNCH(k_density_flat, true, n, sp, sv, rho_p, start, end, n, k, rho0, m, st.common, st.kernels, masks);
    else if (use_fast() && masks)
        FG_SPH_LAUNCH(k_density_fast, true, n, sp, sv, rho_p, start, end, n, k, rho0, m, st.common, st.kernels, masks);
    else if (use_fast())
        FG_SPH_LAUNCH(k_density_fast, false, n, sp, (const float4*)nullptr, rho_p, start, end, n, k, rho0, m, st.common, st.kernels,
                      (unsigned long long*)nullptr);
    else if (st.key_mode == FG_KEY_LINEAR)
        hipLaunchKernelGGL(k_density_ref<FG_KEY_LINEAR>, grid, block, 0, st.stream, sp, rho_p, start, end, n, k,
                           rho0, m, st.common, st.kernels);
    else
        hipLaunchKernelGGL(k_density_ref<FG_KEY_REF_XOR>, grid, block, 0, st.stream, sp, rho_p, start, end, n, k,
                           rho0, m, st.common, st.kernels);
    return check_launch("density") ? FG_OK : fg_peek_error();
}

int launch_forces(float4* sp, float4* sv, float2* rho_p, uint32_t* start, uint32_t* end, float4* vout,
                  float4* pout, uint32_t* id, uint32_t n, float m, float mu, const unsigned long long* masks = nullptr,
                  const Contacts* bcp = nullptr) {
    if (!stencil_covers_h("force pass")) return FG_ERR_UNSUPPORTED;
    Contacts bc{};
    if (bcp) bc = *bcp;
    const bool coupled = bc.coupled != 0;
    const State& st = state();
    const dim3 grid(grid_for(n, kSphThreads)), block(kSphThreads);
    if (use_fast() && masks && nbr_format_tiled(masks)) {
        if (coupled)
            hipLaunchKernelGGL(k_forces_tile<true>, dim3(grid_for(n, kTileLanes)), dim3(kTileLanes), 0, st.stream, sp,
                               sv, rho_p, start, end, vout, pout, id, n, m, mu, st.common, st.kernels, masks, bc);
        else
            hipLaunchKernelGGL(k_forces_tile<false>, dim3(grid_for(n, kTileLanes)), dim3(kTileLanes), 0, st.stream,
                               sp, sv, rho_p, start, end, vout, pout, id, n, m, mu, st.common, st.kernels, masks, bc);
    } else if (use_fast() && masks && st.flat_forces) {
        if (coupled)
            FG_SPH_LAUNCH(k_forces_flat, true, n, sp, sv, rho_p, start, end, vout, pout, id, n, m, mu, st.common,
                          st.kernels, masks, bc);
        else
            FG_SPH_LAUNCH(k_forces_flat, false, n, sp, sv, rho_p, start, end, vout, pout, id, n, m, mu, st.common,
                          st.kernels, masks, bc);
    } else if (use_fast()) {
        const unsigned long long* mk = masks;
        if (coupled && mk && st.defer_contact) {
            // The deferred particles' pair sums and list live in the neighbour
            // buffer's list area (used otherwise only by the tiled passes) and
            // the count in its count area; the count is cleared by a kernel,
            // not a memset, so the step can be captured in a graph.
            char* base = reinterpret_cast<char*>(const_cast<unsigned long long*>(mk));
            DeferredContact dc{reinterpret_cast<float4*>(base + nbr_list_offset(n)),
                               reinterpret_cast<uint32_t*>(base + nbr_list_offset(n) + 16 * (size_t)n),
                               reinterpret_cast<uint32_t*>(base + nbr_count_offset(n))};
            launch_fill_u32(dc.count, 0u, 1, st.stream);
            switch (st.sph_block) {
                case 64: FG_SPH_LAUNCH_BS(k_forces_deferred<64>, 64, n, sp, sv, rho_p, start, end, vout, pout, id, n,
                                          m, mu, st.common, st.kernels, mk, bc, dc); break;
                case 128: FG_SPH_LAUNCH_BS(k_forces_deferred<128>, 128, n, sp, sv, rho_p, start, end, vout, pout, id,
                                           n, m, mu, st.common, st.kernels, mk, bc, dc); break;
                case 512: FG_SPH_LAUNCH_BS(k_forces_deferred<512>, 512, n, sp, sv, rho_p, start, end, vout, pout, id,
                                           n, m, mu, st.common, st.kernels, mk, bc, dc); break;
                default: FG_SPH_LAUNCH_BS(k_forces_deferred<256>, 256, n, sp, sv, rho_p, start, end, vout, pout, id,
                                          n, m, mu, st.common, st.kernels, mk, bc, dc);
            }
            // up to one lane pair per particle (the count is known only on the device)
            unsigned blocks = grid_for((uint64_t)n * kFinishLanes, 256);
            if (blocks > 8192) blocks = 8192;
            hipLaunchKernelGGL(k_contact_finish, dim3(blocks), dim3(256), 0, st.stream, sp, sv, vout, pout, id, n, m,
                               st.common, bc, dc);
        } else if (coupled && mk)
            FG_SPH_LAUNCH2(k_forces_fast, true, true, n, sp, sv, rho_p, start, end, vout, pout, id, n, m, mu,
                           st.common, st.kernels, mk, bc);
        else if (coupled)
            FG_SPH_LAUNCH2(k_forces_fast, false, true, n, sp, sv, rho_p, start, end, vout, pout, id, n, m, mu,
                           st.common, st.kernels, mk, bc);
        else if (mk)
            FG_SPH_LAUNCH2(k_forces_fast, true, false, n, sp, sv, rho_p, start, end, vout, pout, id, n, m, mu,
                           st.common, st.kernels, mk, bc);
        else
            FG_SPH_LAUNCH2(k_forces_fast, false, false, n, sp, sv, rho_p, start, end, vout, pout, id, n, m, mu,
                           st.common, st.kernels, mk, bc);
    } else if (st.key_mode == FG_KEY_LINEAR) {
        if (coupled)
            hipLaunchKernelGGL((k_forces_ref<FG_KEY_LINEAR, true>), grid, block, 0, st.stream, sp, sv, rho_p, start,
                               end, vout, pout, id, n, m, mu, st.common, st.kernels, bc);
        else
            hipLaunchKernelGGL((k_forces_ref<FG_KEY_LINEAR, false>), grid, block, 0, st.stream, sp, sv, rho_p, start,
                               end, vout, pout, id, n, m, mu, st.common, st.kernels, bc);
    } else {
        if (coupled)
            hipLaunchKernelGGL((k_forces_ref<FG_KEY_REF_XOR, true>), grid, block, 0, st.stream, sp, sv, rho_p, start,
                               end, vout, pout, id, n, m, mu, st.common, st.kernels, bc);
        else
            hipLaunchKernelGGL((k_forces_ref<FG_KEY_REF_XOR, false>), grid, block, 0, st.stream, sp, sv, rho_p,
                               start, end, vout, pout, id, n, m, mu, st.common, st.kernels, bc);
    }
    return check_launch("forces") ? FG_OK : fg_peek_error();
}

}  // namespace fg

using namespace fg;

extern "C" {

// K_FluidSim.cu:488-518
void gCalcMassDensity(fg_float4* orderedPositions, fg_float2* mass_pressure, unsigned int* hashTable,
                      unsigned int* gridStart, unsigned int* gridEnd, int numParticles, float gasConstant,
                      float restDensity, float mass) {
    (void)hashTable;
    if (numParticles <= 0) return;
    if (!have_params("gCalcMassDensity")) return;
    if (!orderedPositions || !mass_pressure || !gridStart || !gridEnd) {
        set_error(FG_ERR_INVALID_ARGUMENT, "gCalcMassDensity: null pointer");
        return;
    }
    launch_density(FG_F4(orderedPositions), FG_F2(mass_pressure), gridStart, gridEnd, (uint32_t)numParticles,
                   gasConstant, restDensity, mass);
}

// K_FluidSim.cu:527-566
void gSPHCalculateForces(fg_float4* orderedPositions, fg_float4* orderedVelocity, fg_float4* forces,
                         fg_float2* density_pressure, unsigned int* gridStart, unsigned int* gridEnd, fg_float4* vel,
                         unsigned int* particleId, unsigned int numParticles, float mass, float viscosity) {
    (void)forces;
    if (numParticles == 0) return;
    if (!have_params("gSPHCalculateForces")) return;
    if (!orderedPositions || !orderedVelocity || !density_pressure || !gridStart || !gridEnd || !vel ||
        !particleId) {
        set_error(FG_ERR_INVALID_ARGUMENT, "gSPHCalculateForces: null pointer");
        return;
    }
    if ((void*)vel == (void*)orderedVelocity) {
        set_error(FG_ERR_INVALID_ARGUMENT, "gSPHCalculateForces: vel must not alias orderedVelocity");
        return;
    }
    launch_forces(FG_F4(orderedPositions), FG_F4(orderedVelocity), FG_F2(density_pressure), gridStart, gridEnd,
                  FG_F4(vel), nullptr, particleId, numParticles, mass, viscosity);
}

// K_FluidSim.cu:571-586
void gSPHIntegrate(fg_float4* _in_sph_positions, fg_float4* _in_sph_velocity, fg_float4* _in_sph_forces,
                   unsigned int _in_num_particles, fg_float2* _in_density_press, unsigned int* _in_grid_start,
                   unsigned int* _in_grid_end) {
    (void)_in_sph_forces;
    (void)_in_density_press;
    (void)_in_grid_start;
    (void)_in_grid_end;
    if (_in_num_particles == 0) return;
    if (!have_params("gSPHIntegrate")) return;
    if (!_in_sph_positions || !_in_sph_velocity) {
        set_error(FG_ERR_INVALID_ARGUMENT, "gSPHIntegrate: null pointer");
        return;
    }
    const State& st = state();
    float q;
    {
#pragma clang fp contract(off)
        q = st.common.deltaTimeIntegration / st.common.simulation_scale;
    }
    hipLaunchKernelGGL(k_integrate, dim3(grid_for(_in_num_particles, 256)), dim3(256), 0, st.stream,
                       FG_F4(_in_sph_positions), FG_F4(_in_sph_velocity), _in_num_particles, q);
    check_launch("gSPHIntegrate");
}

int fg_sph_forces_integrate(fg_float4* orderedPositions, fg_float4* orderedVelocity, fg_float2* density_pressure,
                            unsigned int* gridStart, unsigned int* gridEnd, fg_float4* vel, fg_float4* pos,
                            unsigned int* particleId, unsigned int numParticles, float mass, float viscosity) {
    if (numParticles == 0) return FG_OK;
    if (!have_params("fg_sph_forces_integrate")) return FG_ERR_NOT_INITIALIZED;
    if (!orderedPositions || !orderedVelocity || !density_pressure || !gridStart || !gridEnd || !vel || !pos ||
        !particleId) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sph_forces_integrate: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    return launch_forces(FG_F4(orderedPositions), FG_F4(orderedVelocity), FG_F2(density_pressure), gridStart, gridEnd,
                         FG_F4(vel), FG_F4(pos), particleId, numParticles, mass, viscosity);
}

size_t fg_neighbor_mask_bytes(unsigned int numParticles) {
    // 9 row masks, then (each from a 256-byte boundary) a 32-byte neighbour
    // record, a u16 list count and a list of kListMax u16 per particle
    return nbr_buffer_bytes(numParticles);
}

int fg_sph_density_nbr(fg_float4* orderedPositions, const fg_float4* orderedVelocity, fg_float2* mass_pressure,
                       unsigned int* gridStart,
                       unsigned int* gridEnd, unsigned long long* neighborMasks, unsigned int numParticles,
                       float gasConstant, float restDensity, float mass) {
    if (numParticles == 0) return FG_OK;
    if (!have_params("fg_sph_density_nbr")) return FG_ERR_NOT_INITIALIZED;
    if (!orderedPositions || !orderedVelocity || !mass_pressure || !gridStart || !gridEnd || !neighborMasks) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sph_density_nbr: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    return launch_density(FG_F4(orderedPositions), FG_F2(mass_pressure), gridStart, gridEnd, numParticles,
                          gasConstant, restDensity, mass, neighborMasks,
                          reinterpret_cast<const float4*>(orderedVelocity));
}

int fg_sph_forces_integrate_nbr(fg_float4* orderedPositions, fg_float4* orderedVelocity,
                                fg_float2* density_pressure, unsigned int* gridStart, unsigned int* gridEnd,
                                const unsigned long long* neighborMasks, fg_float4* vel, fg_float4* pos,
                                unsigned int* particleId, unsigned int numParticles, float mass, float viscosity) {
    if (numParticles == 0) return FG_OK;
    if (!have_params("fg_sph_forces_integrate_nbr")) return FG_ERR_NOT_INITIALIZED;
    if (!orderedPositions || !orderedVelocity || !density_pressure || !gridStart || !gridEnd || !neighborMasks ||
        !vel || !particleId) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sph_forces_integrate_nbr: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    return launch_forces(FG_F4(orderedPositions), FG_F4(orderedVelocity), FG_F2(density_pressure), gridStart, gridEnd,
                         FG_F4(vel), pos ? FG_F4(pos) : nullptr, particleId, numParticles, mass, viscosity,
                         neighborMasks);
}

}  // extern "C"
