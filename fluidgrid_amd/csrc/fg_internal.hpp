// fg_internal.hpp -- shared host state, error handling and device helpers of
// libfluidgrid_amd.so.  Device code here is written for gfx950 (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/fluidgrid_amd.h"

namespace fg {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;

// ---------------------------------------------------------------------------
// Host-side state (the reference keeps parameters in __constant__ memory,
// K_FluidSim.cu:37-40; here they are held on the host and passed by value as
// kernel arguments, which the hardware reads through the scalar cache).
// ---------------------------------------------------------------------------
struct State {
    fg_common_fluid_params common{};
    fg_kernels_cache kernels{};
    bool have_common = false;
    bool have_kernels = false;
    int key_mode = FG_KEY_REF_XOR;
    int math_mode = FG_MATH_FAST;
    hipStream_t stream = nullptr;
    // Launch shape chosen by the tuner (fg_set_tuning): workgroup size of the
    // fast density/force kernels, and digit bits per radix-sort pass.  64-lane
    // workgroups measured fastest for the LDS-staged passes (1M, 8M and 33.5M
    // particles: 2.6%, 3.4% and 1.8% shorter steps than 256).
    int sph_block = 64;
    int sort_digit_bits = 8;
    // Schedules of the masked fast path: flat candidate/neighbour stream per
    // lane (true) or row by row (false); results are identical.
    bool flat_forces = false;
    bool flat_density = false;
    // LDS-tiled single-fluid passes (density lists + list-walking forces);
    // they take precedence over the two schedules above when switched on.
    bool tiled = false;
    // Coupled masked force pass (row loop): body contact of the particles
    // inside a body grid's bounds deferred to a second kernel that walks each
    // one's body-grid cells on a pair of lanes (fg_set_contact_defer);
    // results are identical.
    bool defer_contact = true;
    // Arithmetic of the fused density pass (fg_set_precision): FP32, or FP16
    // culling with exact FP32 survivors; the latter takes precedence over the
    // schedules above for density.
    int precision = FG_PRECISION_FP32;
};

State& state();

// Records that a cell pass wrote the FP16 pair records of n particles into a
// neighbour buffer (the mixed density pass requires it).
void note_half_records(const void* buf, uint32_t n);

// Thread-local error slot.
void set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
bool check_hip(hipError_t e, const char* what);
// Records the launch status of the last kernel; returns true on success.
bool check_launch(const char* what);

// Fill kernel launcher (graph-safe replacement for hipMemsetAsync).
void launch_fill_u32(uint32_t* dst, uint32_t value, uint64_t count, hipStream_t s);

// Geometry of one bucket grid: the fluid grid, or a body-sphere grid.
struct GridDesc {
    float3 vmin;       // LINEAR origin
    float3 cs;         // cell size
    int3 dims;         // LINEAR cells per axis
    uint32_t nb;       // number of buckets (key range)
    int mode;          // FG_KEY_REF_XOR or FG_KEY_LINEAR
    uint32_t key_base; // first bucket of this grid in a shared key space (batched scenes); 0 otherwise
};

// Penalty contact of fluid particles against a body-sphere grid
// (K_DeviceFunctions.cu:32-159 as called from K_FluidSim.cu:380-386).
struct BodyContact {
    const float4* pos;      // body spheres in grid order, w = body index
    const float4* vel;      // per-body velocity, indexed by w
    const uint32_t* start;  // body grid cell ranges
    const uint32_t* end;
    const float4* aabb;     // [2] expanded bounds of all spheres, or null
    GridDesc grid;          // REF_XOR body grid
    int3 relation;          // stencil half-width per axis
    float fluid_radius, body_radius, ks, kd, eps;
    uint32_t count;         // number of body spheres
    uint32_t nvel;          // number of entries of vel
    int enabled;
};

// Body grids seen by the fluid force pass; `coupled` selects the force
// expression of K_FluidSim.cu:389-390 (external term added after the mass
// factor) over that of K_FluidSim.cu:287.
struct Contacts {
    BodyContact st, dyn;
    int coupled;
};

// Deferred body contact: the force pass hands particles on which contact
// can act to a second pass (contact, walls, clamp, integration) with their
// pair sum `internal`, appended to a list; the rest it finishes itself.
struct DeferredContact {
    float4* acc;      // internal acceleration per sorted particle (xyz)
    uint32_t* list;   // sorted indices of the deferred particles
    uint32_t* count;  // their number
};

// Rigid bodies of one scene of a batch.  All scenes' bodies share global
// arrays (bodies scene-major, spheres in body order, w = global body index);
// each scene's body grid occupies its own bucket range (grid.key_base) of
// one shared body key space, as the scenes' fluid grids do, so one stable
// sort and one cell-range pass serve every scene and every walk stays in
// its scene.  Each field is what the single-fluid RigidBodySimulator
// derives from that scene's parameter blocks.
struct BatchBodyScene {
    BodyContact dyn;        // the body grid as the scene's fluid sees it (aabb: the scene's sphere bounds)
    GridDesc fluid;         // the scene's fluid grid (LINEAR, key_base = the scene's fluid key base)
    int3 rb_to_fluid;       // stencil of the body spheres in the fluid grid
    float sphere_radius;    // body sphere radius
    float fluid_radius;     // 2 * pointRadius of the scene's fluid
    float ks, kd;           // -springForce, -dampingForce
    float bounds_margin;    // (2 pointRadius + r + 0.003) * 1.01: sphere bounds of the fluid-side cull
    float pair_margin;      // (2 r + 0.001) * 1.01: per-body bounds of the body-body broad phase
    uint32_t body_begin, body_end;      // global body range
    uint32_t sphere_begin, sphere_end;  // global sphere range
    // body integration (IntegrateRBK parameters of the scene)
    float dt, vlimit;
    float4 ext;
    float3 vmin, vmax;
};

inline uint32_t num_buckets(const fg_common_fluid_params& p) {
    return (uint32_t)p.bucket_size.x * (uint32_t)p.bucket_size.y * (uint32_t)p.bucket_size.z;
}

inline unsigned grid_for(uint64_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

// float -> int with truncation, saturation and NaN -> 0, spelled out so the
// result does not depend on how the compiler lowers an out-of-range cast.
__device__ __forceinline__ int32_t f2i_sat(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}

// Correctly rounded IEEE single operations for the reference-arithmetic
// code.  They are written as plain operators in a region with contraction
// off, so no multiply is ever fused into a following add, even after
// inlining into code compiled with contraction on (the HIP header's
// __fadd_rn/__fmul_rn are plain operators compiled with contraction
// allowed, and its __fsqrt_rn is the hardware approximation).  Division and
// __builtin_sqrtf are correctly rounded under HIP's default single-precision
// division/square-root mode.
#pragma clang fp contract(off)
__device__ __forceinline__ float rn_add(float a, float b) { return a + b; }
__device__ __forceinline__ float rn_sub(float a, float b) { return a - b; }
__device__ __forceinline__ float rn_mul(float a, float b) { return a * b; }
__device__ __forceinline__ float rn_div(float a, float b) { return a / b; }
__device__ __forceinline__ float rn_sqrt(float a) { return __builtin_sqrtf(a); }
#pragma clang fp contract(fast)

// REF_XOR key: K_DeviceFunctions.cu:17-26.  The add and divide are IEEE
// single operations, so keys are bit-identical to a host evaluation of the
// same expression.
__device__ __forceinline__ uint32_t hash_ref_xor(float px, float py, float pz, float csx, float csy,
                                                 float csz, uint32_t table) {
    int32_t x = f2i_sat(rn_div(rn_add(px, 5000.0f), csx));
    int32_t y = f2i_sat(rn_div(rn_add(py, 5000.0f), csy));
    int32_t z = f2i_sat(rn_div(rn_add(pz, 5000.0f), csz));
    uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
    return h % table;
}

// LINEAR cell coordinate: clamp(floor((p - vmin)/cs), 0, n-1).  Clamping in
// float before the conversion gives the same answer as converting with
// saturation first (NaN maps to 0 in both).
__device__ __forceinline__ int cell_coord(float p, float vmin, float cs, int n) {
    float q = floorf(rn_div(rn_sub(p, vmin), cs));
    q = fminf(fmaxf(q, 0.0f), (float)(n - 1));
    return (int)q;
}

struct CellXYZ { int x, y, z; };

// Stencil reach along one axis: the number of cells on each side of a
// particle's own cell that can hold a neighbour within h, ceil(h / cs) (h in
// world units, kernel_radius / simulation_scale), at least 1.  The 27-cell
// stencil of the density and force passes (the reference's, and the rows of
// the fast path) needs 1: cells at least h on every axis.
__host__ __device__ inline int linear_reach(float kernel_radius, float simulation_scale, float cell) {
    const float q = (kernel_radius / simulation_scale) / cell;
    const int r = (int)ceilf(q);
    return r < 1 ? 1 : r;
}

__device__ __forceinline__ CellXYZ grid_cell(float px, float py, float pz, const GridDesc& g) {
    return {cell_coord(px, g.vmin.x, g.cs.x, g.dims.x), cell_coord(py, g.vmin.y, g.cs.y, g.dims.y),
            cell_coord(pz, g.vmin.z, g.cs.z, g.dims.z)};
}

__device__ __forceinline__ uint32_t grid_linear_key(int cx, int cy, int cz, const GridDesc& g) {
    return (uint32_t)cx + (uint32_t)g.dims.x * ((uint32_t)cy + (uint32_t)g.dims.y * (uint32_t)cz);
}

// Key of a position in either mode.
__device__ __forceinline__ uint32_t grid_key(float px, float py, float pz, const GridDesc& g) {
    if (g.mode == FG_KEY_REF_XOR) return g.key_base + hash_ref_xor(px, py, pz, g.cs.x, g.cs.y, g.cs.z, g.nb);
    CellXYZ c = grid_cell(px, py, pz, g);
    return g.key_base + grid_linear_key(c.x, c.y, c.z, g);
}

__device__ __forceinline__ CellXYZ linear_cell(float px, float py, float pz, const fg_common_fluid_params& cp) {
    return {cell_coord(px, cp.volume_min.x, cp.cell_size.x, cp.bucket_size.x),
            cell_coord(py, cp.volume_min.y, cp.cell_size.y, cp.bucket_size.y),
            cell_coord(pz, cp.volume_min.z, cp.cell_size.z, cp.bucket_size.z)};
}

__device__ __forceinline__ uint32_t linear_key(int cx, int cy, int cz, const fg_common_fluid_params& cp) {
    return (uint32_t)cx + (uint32_t)cp.bucket_size.x * ((uint32_t)cy + (uint32_t)cp.bucket_size.y * (uint32_t)cz);
}

__device__ __forceinline__ uint32_t dev_num_buckets(const fg_common_fluid_params& cp) {
    return (uint32_t)cp.bucket_size.x * (uint32_t)cp.bucket_size.y * (uint32_t)cp.bucket_size.z;
}

// Small float3 helpers with explicit rounding for the reference-arithmetic
// kernels.  Contraction is switched off for this region so that no multiply
// and add below is fused after inlining.
#pragma clang fp contract(off)
struct F3 { float x, y, z; };
__device__ __forceinline__ F3 f3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ F3 add(F3 a, F3 b) { return {rn_add(a.x, b.x), rn_add(a.y, b.y), rn_add(a.z, b.z)}; }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return {rn_sub(a.x, b.x), rn_sub(a.y, b.y), rn_sub(a.z, b.z)}; }
__device__ __forceinline__ F3 mul(F3 a, float s) { return {rn_mul(a.x, s), rn_mul(a.y, s), rn_mul(a.z, s)}; }
__device__ __forceinline__ F3 mul(float s, F3 a) { return {rn_mul(s, a.x), rn_mul(s, a.y), rn_mul(s, a.z)}; }
__device__ __forceinline__ F3 divs(F3 a, float s) { return {rn_div(a.x, s), rn_div(a.y, s), rn_div(a.z, s)}; }
__device__ __forceinline__ float dot(F3 a, F3 b) {
    return rn_add(rn_add(rn_mul(a.x, b.x), rn_mul(a.y, b.y)), rn_mul(a.z, b.z));
}
__device__ __forceinline__ float len(F3 a) { return rn_sqrt(dot(a, a)); }
__device__ __forceinline__ F3 xyz(float4 p) { return {p.x, p.y, p.z}; }
#pragma clang fp contract(fast)

}  // namespace fg

// Reinterpretation between the public POD vector types and HIP's.
static_assert(sizeof(fg_float4) == sizeof(float4) && alignof(fg_float4) == alignof(float4), "float4 layout");
static_assert(sizeof(fg_float2) == sizeof(float2) && alignof(fg_float2) == alignof(float2), "float2 layout");
static_assert(sizeof(fg_float3) == sizeof(float3), "float3 layout");
static_assert(sizeof(fg_common_fluid_params) == 104, "CommonFluidParams layout");
static_assert(sizeof(fg_common_rigid_body_params) == 112, "CommonRigidBodyParams layout");
static_assert(sizeof(fg_kernels_cache) == 12, "KernelsCache layout");
static_assert(sizeof(fg_fluid_params) == 28, "FluidParams layout");
static_assert(sizeof(fg_inflow_desc) == 72, "fg_inflow_desc layout");
static_assert(sizeof(fg_scene_desc) == 160, "scene descriptor layout");

#define FG_F4(p) reinterpret_cast<float4*>(p)
#define FG_F2(p) reinterpret_cast<float2*>(p)
