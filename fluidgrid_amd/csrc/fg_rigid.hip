// fg_rigid.hip -- two-way fluid/rigid-body coupling (the work of
// GPU_Kernels/K_RigidBodySim.cu and the coupled branch of K_FluidSim.cu):
//  * fluid forces with penalty contact against static and dynamic body grids;
//  * per body sphere: penalty force from the fluid, from the spheres of the
//    other bodies and from the static spheres, and torque about the body
//    centre of mass;
//  * per-body force/torque sums in FP64 (replacing the CUDPP FP32 segmented
//    scan), one workgroup per body with a fixed reduction order;
//  * body centre of mass and inertia tensor in FP64 from the sphere set;
//  * body integration (IntegrateRBK) and the reference's reduction helpers.
#include "fg_batch_bodies.hpp"
#include "fg_contact.hpp"
#include "fg_internal.hpp"
#include "fg_quat.hpp"

namespace fg {

int launch_forces(float4* sp, float4* sv, float2* rho_p, uint32_t* start, uint32_t* end, float4* vout, float4* pout,
                  uint32_t* id, uint32_t n, float m, float mu, const unsigned long long* masks, const Contacts* bcp);
GridDesc fluid_grid_desc(const fg_common_fluid_params& cp, fg_float3 cs, uint32_t nb, int mode);

#pragma clang fp contract(off)

// Lanes per body sphere of the body-force kernels: the stencil cells of a
// sphere are walked by a group of kBodyLanes lanes (contact_stencil_lanes),
// so a sphere's 27 chains of dependent loads are in flight together.
constexpr uint32_t kBodyLanes = 8;

// K_RigidBodySim.cu:66-99, for body sphere i, with its stencil walked by LP
// lanes (sub: the lane's index in its group; LP = 1: one lane, the 27 cell
// starts loaded together).
template <uint32_t LP>
__device__ __forceinline__ void body_sphere_from_fluid(uint32_t i, const float4* __restrict__ rb_pos,
                                                       const float4* __restrict__ rb_vel,
                                                       const float4* __restrict__ rb_cm, float4* __restrict__ out_f,
                                                       float4* __restrict__ out_t, uint32_t nbodies, float radius,
                                                       const float4* __restrict__ sph_pos,
                                                       const float4* __restrict__ sph_vel,
                                                       const uint32_t* __restrict__ sph_start,
                                                       const uint32_t* __restrict__ sph_end, uint32_t nsph,
                                                       float sph_radius, const GridDesc& g, int3 relation, float ks,
                                                       float kd, uint32_t sub = 0) {
    const float4 pos = rb_pos[i];
    const int b = f2i_sat(pos.w);
    const bool ok = b >= 0 && (uint32_t)b < nbodies;
    const float4 vel = ok ? rb_vel[b] : make_float4(0, 0, 0, 0);
    const float4 cm = ok ? rb_cm[b] : make_float4(0, 0, 0, 0);
    F3 F;
    if constexpr (LP == 1)
        F = contact_stencil<true>(pos, vel, radius, sph_pos, sph_vel, sph_radius, ks, kd, relation, sph_start, sph_end,
                                  g, false, false, 0, 0.0001f, nsph, nsph);
    else
        F = contact_stencil_lanes<LP>(pos, vel, radius, sph_pos, sph_vel, sph_radius, ks, kd, relation, sph_start,
                                      sph_end, g, false, false, 0, 0.0001f, nsph, nsph, sub);
    if (sub != 0) return;
    out_f[i] = make_float4(F.x, F.y, F.z, (float)b);
    const float4 dir = make_float4(pos.x - cm.x, pos.y - cm.y, pos.z - cm.z, pos.w - cm.w);
    float4 t = vec3_cross(dir, make_float4(F.x, F.y, F.z, 0.0f));
    t.w = (float)b;
    out_t[i] = t;
}

__global__ void __launch_bounds__(256)
    k_body_forces_from_fluid(const float4* __restrict__ rb_pos, const float4* __restrict__ rb_vel,
                             const float4* __restrict__ rb_cm, float4* __restrict__ out_f, float4* __restrict__ out_t,
                             uint32_t m, uint32_t nbodies, float radius, const float4* __restrict__ sph_pos,
                             const float4* __restrict__ sph_vel, const uint32_t* __restrict__ sph_start,
                             const uint32_t* __restrict__ sph_end, uint32_t nsph, float sph_radius, GridDesc g,
                             int3 relation, float ks, float kd) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = t / kBodyLanes;
    if (i >= m) return;  // the same for every lane of a group
    body_sphere_from_fluid<kBodyLanes>(i, rb_pos, rb_vel, rb_cm, out_f, out_t, nbodies, radius, sph_pos, sph_vel,
                                       sph_start, sph_end, nsph, sph_radius, g, relation, ks, kd, t % kBodyLanes);
}

__device__ __forceinline__ bool inside_box(const float4& p, const float4& lo, const float4& hi) {
    return p.x >= lo.x && p.y >= lo.y && p.z >= lo.z && p.x <= hi.x && p.y <= hi.y && p.z <= hi.z;
}

// K_RigidBodySim.cu:18-61: per body sphere, penalty contact against the
// spheres of the other dynamic bodies (those of its own body are skipped by
// w) and against the static spheres; torque about the body's centre of mass.
// The broad phase replaces the host one of RigidBodySimulator.cpp:334-406:
// with body bounds (per body, the box of its sphere centres expanded by the
// contact distance) the dynamic walk is skipped for a sphere inside no other
// body's box, and the static walk for a sphere outside the static box; a
// skipped walk would only have added +0.  ACCUM adds the force and torque to
// what out_f/out_t hold (the fluid's part), else they are written.
// For body sphere i; the broad phase looks at the bounds of bodies
// [k0, k1) other than the sphere's own.  The stencils are walked by LP
// lanes (sub: the lane's index in its group).
template <bool ACCUM, uint32_t LP>
__device__ __forceinline__ void body_sphere_from_bodies(uint32_t i, const float4* __restrict__ rb_pos,
                                                        const float4* __restrict__ rb_vel,
                                                        const float4* __restrict__ rb_cm, float4* __restrict__ out_f,
                                                        float4* __restrict__ out_t, uint32_t nbodies, float radius,
                                                        const BodyContact& dyn, const BodyContact& st,
                                                        const float4* __restrict__ body_bounds, uint32_t k0,
                                                        uint32_t k1, uint32_t sub) {
    const float4 pos = rb_pos[i];
    const int b = f2i_sat(pos.w);
    const bool ok = b >= 0 && (uint32_t)b < nbodies;
    const float4 vel = ok ? rb_vel[b] : make_float4(0, 0, 0, 0);
    const float4 cm = ok ? rb_cm[b] : make_float4(0, 0, 0, 0);
    F3 F = f3(0.0f, 0.0f, 0.0f);
    if (dyn.enabled) {
        bool near = body_bounds == nullptr;
        if (!near) {
            // the group's lanes test the bodies in turn (lane sub: k0 + sub, k0 + sub + LP, ...)
            bool hit = false;
            for (uint32_t k = k0 + sub; !hit && k < k1; k += LP)
                if ((int)k != b) hit = inside_box(pos, body_bounds[2 * k], body_bounds[2 * k + 1]);
            const uint32_t first = __lane_id() & ~(LP - 1u);
            near = ((__ballot(hit) >> first) & (LP == 32 ? 0xFFFFFFFFull : ((1ull << LP) - 1ull))) != 0ull;
        }
        if (near)
            F = add(F, contact_stencil_lanes<LP>(pos, vel, radius, dyn.pos, dyn.vel, dyn.body_radius, dyn.ks, dyn.kd,
                                                 dyn.relation, dyn.start, dyn.end, dyn.grid, true, true, b, dyn.eps,
                                                 dyn.count, dyn.nvel, sub));
    }
    if (st.enabled && (st.aabb == nullptr || inside_box(pos, st.aabb[0], st.aabb[1])))
        F = add(F, contact_stencil_lanes<LP>(pos, vel, radius, st.pos, nullptr, st.body_radius, st.ks, st.kd,
                                             st.relation, st.start, st.end, st.grid, false, false, 0, st.eps,
                                             st.count, 0u, sub));
    if (sub != 0) return;
    const float4 dir = make_float4(pos.x - cm.x, pos.y - cm.y, pos.z - cm.z, pos.w - cm.w);
    const float4 t = vec3_cross(dir, make_float4(F.x, F.y, F.z, 0.0f));
    if (ACCUM) {
        const float4 f0 = out_f[i], t0 = out_t[i];
        out_f[i] = make_float4(f0.x + F.x, f0.y + F.y, f0.z + F.z, (float)b);
        out_t[i] = make_float4(t0.x + t.x, t0.y + t.y, t0.z + t.z, (float)b);
    } else {
        out_f[i] = make_float4(F.x, F.y, F.z, (float)b);
        out_t[i] = make_float4(t.x, t.y, t.z, (float)b);
    }
}

template <bool ACCUM>
__global__ void __launch_bounds__(256)
    k_body_forces_from_bodies(const float4* __restrict__ rb_pos, const float4* __restrict__ rb_vel,
                              const float4* __restrict__ rb_cm, float4* __restrict__ out_f,
                              float4* __restrict__ out_t, uint32_t m, uint32_t nbodies, float radius, BodyContact dyn,
                              BodyContact st, const float4* __restrict__ body_bounds, uint32_t nbounds) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = t / kBodyLanes;
    if (i >= m) return;  // the same for every lane of a group
    body_sphere_from_bodies<ACCUM, kBodyLanes>(i, rb_pos, rb_vel, rb_cm, out_f, out_t, nbodies, radius, dyn, st,
                                               body_bounds, 0u, nbounds, t % kBodyLanes);
}

// K_RigidBodySim.cu:102-111
__global__ void k_copy_rb_data(uint32_t* __restrict__ out_rb, uint32_t* __restrict__ out_idx,
                               const float4* __restrict__ f, uint64_t m) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    out_rb[i] = (uint32_t)f[i].w;
    out_idx[i] = (uint32_t)i;
}

// K_RigidBodySim.cu:114-134
__global__ void k_reorder_force_torque(float4* __restrict__ of, float4* __restrict__ ot, const float4* __restrict__ f,
                                       const float4* __restrict__ t, const uint32_t* __restrict__ idx, uint32_t m) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t s = idx[i];
    if (s >= m) return;
    of[i] = f[s];
    ot[i] = t[s];
}

// K_RigidBodySim.cu:140-155 (one segment per body, so the += is race-free)
__global__ void k_add_force_torque(const float4* __restrict__ sum_f, const float4* __restrict__ sum_t,
                                   const float4* __restrict__ ordered_f, float4* __restrict__ to_f,
                                   float4* __restrict__ to_t, const int* __restrict__ offset, uint64_t nb) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const int o = offset[i];
    const int b = (int)ordered_f[o].w;
    to_f[b] = quat_add(to_f[b], sum_f[o]);
    to_t[b] = quat_add(to_t[b], sum_t[o]);
}

// K_RigidBodySim.cu:160-262, for body idx.
__device__ __forceinline__ void integrate_body(uint64_t idx, float4* __restrict__ pos, const float* __restrict__ mass,
                                               float4* __restrict__ vel, const float4* __restrict__ forces, float dt,
                                               float4 ext, float3 vmin, float3 vmax, float4* __restrict__ L,
                                               const float4* __restrict__ torque,
                                               const float4* __restrict__ inv_inertia, float4* __restrict__ orient,
                                               float vlimit) {
    float4 F = forces[idx];
    const float spd = sqrtf(F.x * F.x + F.y * F.y + F.z * F.z + F.w * F.w);
    if (spd > vlimit) F = f4_scale(F, vlimit / spd);
    float4 p = pos[idx];
    const float mm = mass[idx];
    const float4 acc = make_float4(((F.x + ext.x) / mm) * dt, ((F.y + ext.y) / mm) * dt, ((F.z + ext.z) / mm) * dt,
                                   ((F.w + ext.w) / mm) * dt);
    float4 v = quat_add(vel[idx], f4_scale(acc, dt));
    p.x += v.x * dt;
    p.y += v.y * dt;
    p.z += v.z * dt;
    float4 l = quat_add(L[idx], f4_scale(torque[idx], dt));
    l = f4_scale(l, 0.9f);
    L[idx] = l;
    const float4 q = orient[idx];
    float4 w = angular_velocity(inv_inertia + 3 * idx, l, q);
    w.w = 0.0f;
    w = f4_scale(quat_product(w, q), 0.5f);
    const float4 nq = quat_normalize(quat_add(q, quat_product(f4_scale(f4_scale(w, dt), 0.1f), q)));
    orient[idx] = nq;
    if (p.y < vmin.y) { p.y = vmin.y; v.y *= -0.7f; }
    if (p.x <= vmin.x) { p.x = vmin.x; v.x *= -0.7f; }
    else if (p.x >= vmax.x) { p.x = vmax.x; v.x *= -0.7f; }
    if (p.z <= vmin.z) { p.z = vmin.z; v.z *= -0.7f; }
    else if (p.z >= vmax.z) { p.z = vmax.z; v.z *= -0.7f; }
    pos[idx] = p;
    vel[idx] = v;
}

__global__ void k_integrate_rb(float4* __restrict__ pos, const float* __restrict__ mass, float4* __restrict__ vel,
                               const float4* __restrict__ forces, uint64_t nb, float dt, float4 ext, float3 vmin,
                               float3 vmax, float4* __restrict__ L, const float4* __restrict__ torque,
                               const float4* __restrict__ inv_inertia, float4* __restrict__ orient, float vlimit) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb) return;
    integrate_body(idx, pos, mass, vel, forces, dt, ext, vmin, vmax, L, torque, inv_inertia, orient, vlimit);
}

#pragma clang fp contract(fast)

// Per-body FP64 sums of per-sphere forces and torques.  Spheres of body b
// occupy [offset[b], offset[b+1]) of the input arrays (body order).  One
// workgroup per body; each lane sums a strided subset in index order, then a
// fixed tree combines the lanes, so the result does not depend on timing.
constexpr int kReduceThreads = 256;
constexpr int kBodySumThreads = 1024;  // 20K spheres: 20 loads per lane, unrolled

__global__ void __launch_bounds__(kBodySumThreads)
    k_body_reduce_f64(const float4* __restrict__ f, const float4* __restrict__ t, const uint32_t* __restrict__ offset,
                      uint32_t nbodies, double4* __restrict__ out_f, double4* __restrict__ out_t,
                      float4* __restrict__ acc_f, float4* __restrict__ acc_t) {
    __shared__ double s[6][kBodySumThreads];
    const uint32_t b = blockIdx.x;
    if (b >= nbodies) return;
    const uint32_t lo = offset[b], hi = offset[b + 1];
    double a[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 4
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kBodySumThreads) {
        const float4 fi = f[i], ti = t[i];
        a[0] += fi.x; a[1] += fi.y; a[2] += fi.z;
        a[3] += ti.x; a[4] += ti.y; a[5] += ti.z;
    }
    for (int k = 0; k < 6; k++) s[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = kBodySumThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; k++) s[k][threadIdx.x] += s[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out_f[b] = make_double4(s[0][0], s[1][0], s[2][0], 0.0);
        out_t[b] = make_double4(s[3][0], s[4][0], s[5][0], 0.0);
        if (acc_f) acc_f[b] = make_float4((float)s[0][0], (float)s[1][0], (float)s[2][0], 0.0f);
        if (acc_t) acc_t[b] = make_float4((float)s[3][0], (float)s[4][0], (float)s[5][0], 0.0f);
    }
}

// Centre of mass (pass 0) or inertia tensor about `cm` (pass 1) in FP64 of
// equal point masses M_b / count_b at the sphere centres.
__global__ void __launch_bounds__(kReduceThreads)
    k_body_mass_props_f64(const float4* __restrict__ world, const uint32_t* __restrict__ offset, uint32_t nbodies,
                          const double* __restrict__ mass, double* __restrict__ cm, double* __restrict__ inertia,
                          int pass) {
    __shared__ double s[6][kReduceThreads];
    const uint32_t b = blockIdx.x;
    if (b >= nbodies) return;
    const uint32_t lo = offset[b], hi = offset[b + 1];
    double a[6] = {0, 0, 0, 0, 0, 0};
    const double c0 = pass ? cm[3 * b] : 0.0, c1 = pass ? cm[3 * b + 1] : 0.0, c2 = pass ? cm[3 * b + 2] : 0.0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kReduceThreads) {
        const double x = world[i].x, y = world[i].y, z = world[i].z;
        if (pass == 0) {
            a[0] += x; a[1] += y; a[2] += z;
        } else {
            const double rx = x - c0, ry = y - c1, rz = z - c2;
            a[0] += ry * ry + rz * rz;  // Ixx
            a[1] += rx * rx + rz * rz;  // Iyy
            a[2] += rx * rx + ry * ry;  // Izz
            a[3] -= rx * ry;            // Ixy
            a[4] -= rx * rz;            // Ixz
            a[5] -= ry * rz;            // Iyz
        }
    }
    for (int k = 0; k < 6; k++) s[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = kReduceThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; k++) s[k][threadIdx.x] += s[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double cnt = (double)(hi - lo);
        if (pass == 0) {
            for (int k = 0; k < 3; k++) cm[3 * b + k] = cnt > 0 ? s[k][0] / cnt : 0.0;
        } else {
            const double mi = cnt > 0 ? mass[b] / cnt : 0.0;
            double* I = inertia + 9 * b;
            I[0] = mi * s[0][0]; I[4] = mi * s[1][0]; I[8] = mi * s[2][0];
            I[1] = I[3] = mi * s[3][0];
            I[2] = I[6] = mi * s[4][0];
            I[5] = I[7] = mi * s[5][0];
        }
    }
}

// Bounds of all sphere centres, expanded by `margin`: out[0] = min, out[1] = max.
__global__ void __launch_bounds__(1024) k_sphere_bounds(const float4* __restrict__ p, uint32_t m, float margin,
                                                        float4* __restrict__ out) {
    __shared__ float s[6][1024];
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (uint32_t i = threadIdx.x; i < m; i += 1024) {
        const float4 q = p[i];
        lo[0] = fminf(lo[0], q.x); lo[1] = fminf(lo[1], q.y); lo[2] = fminf(lo[2], q.z);
        hi[0] = fmaxf(hi[0], q.x); hi[1] = fmaxf(hi[1], q.y); hi[2] = fmaxf(hi[2], q.z);
    }
    for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = lo[k];
        s[3 + k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; k++) {
                s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + w]);
                s[3 + k][threadIdx.x] = fmaxf(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = make_float4(s[0][0] - margin, s[1][0] - margin, s[2][0] - margin, 0.0f);
        out[1] = make_float4(s[3][0] + margin, s[4][0] + margin, s[5][0] + margin, 0.0f);
    }
}

// Per-body bounds: out[2b] = min, out[2b+1] = max of the centres of the
// spheres [offset[b], offset[b+1]), expanded by margin; one workgroup per
// body.  A body without spheres gets an empty box (min > max).
__global__ void __launch_bounds__(256) k_body_bounds(const float4* __restrict__ p, const uint32_t* __restrict__ offset,
                                                     uint32_t nbodies, float margin, float4* __restrict__ out) {
    __shared__ float s[6][256];
    const uint32_t b = blockIdx.x;
    if (b >= nbodies) return;
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (uint32_t i = offset[b] + threadIdx.x; i < offset[b + 1]; i += 256) {
        const float4 q = p[i];
        lo[0] = fminf(lo[0], q.x); lo[1] = fminf(lo[1], q.y); lo[2] = fminf(lo[2], q.z);
        hi[0] = fmaxf(hi[0], q.x); hi[1] = fmaxf(hi[1], q.y); hi[2] = fmaxf(hi[2], q.z);
    }
    for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = lo[k];
        s[3 + k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; k++) {
                s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + w]);
                s[3 + k][threadIdx.x] = fmaxf(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * b] = make_float4(s[0][0] - margin, s[1][0] - margin, s[2][0] - margin, 0.0f);
        out[2 * b + 1] = make_float4(s[3][0] + margin, s[4][0] + margin, s[5][0] + margin, 0.0f);
    }
}

// ---------------------------------------------------------------------------
// Batched scenes: the kernels above, element by element, with the
// parameters of the element's scene (BatchBodyScene).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k_bbody_transform(uint32_t* __restrict__ key, uint32_t* __restrict__ id, const float4* __restrict__ cm,
                      const float4* __restrict__ orient, const float4* __restrict__ rel, float4* __restrict__ world,
                      uint32_t m, const uint32_t* __restrict__ body_scene, const BatchBodyScene* __restrict__ scenes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const float4 p = rel[i];
    const int b = f2i_sat(p.w);
    const float4 a = world_sphere(p, orient[b], cm[b]);
    id[i] = i;
    world[i] = a;
    key[i] = grid_key(a.x, a.y, a.z, scenes[body_scene[b]].dyn.grid);
}

// One workgroup per scene over its spheres (k_sphere_bounds per scene; min
// and max do not depend on the reduction order).
__global__ void __launch_bounds__(256) k_bscene_bounds(const float4* __restrict__ p,
                                                       const BatchBodyScene* __restrict__ scenes, uint32_t nscenes,
                                                       float4* __restrict__ out) {
    __shared__ float s[6][256];
    const uint32_t sc = blockIdx.x;
    if (sc >= nscenes) return;
    const BatchBodyScene& d = scenes[sc];
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (uint32_t i = d.sphere_begin + threadIdx.x; i < d.sphere_end; i += 256) {
        const float4 q = p[i];
        lo[0] = fminf(lo[0], q.x); lo[1] = fminf(lo[1], q.y); lo[2] = fminf(lo[2], q.z);
        hi[0] = fmaxf(hi[0], q.x); hi[1] = fmaxf(hi[1], q.y); hi[2] = fmaxf(hi[2], q.z);
    }
    for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = lo[k];
        s[3 + k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; k++) {
                s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + w]);
                s[3 + k][threadIdx.x] = fmaxf(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float mg = d.bounds_margin;
        out[2 * sc] = make_float4(s[0][0] - mg, s[1][0] - mg, s[2][0] - mg, 0.0f);
        out[2 * sc + 1] = make_float4(s[3][0] + mg, s[4][0] + mg, s[5][0] + mg, 0.0f);
    }
}

// k_body_bounds with the margin of each body's scene.
__global__ void __launch_bounds__(256) k_bbody_bounds(const float4* __restrict__ p, const uint32_t* __restrict__ offset,
                                                      const uint32_t* __restrict__ body_scene,
                                                      const BatchBodyScene* __restrict__ scenes, uint32_t nbodies,
                                                      float4* __restrict__ out) {
    __shared__ float s[6][256];
    const uint32_t b = blockIdx.x;
    if (b >= nbodies) return;
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (uint32_t i = offset[b] + threadIdx.x; i < offset[b + 1]; i += 256) {
        const float4 q = p[i];
        lo[0] = fminf(lo[0], q.x); lo[1] = fminf(lo[1], q.y); lo[2] = fminf(lo[2], q.z);
        hi[0] = fmaxf(hi[0], q.x); hi[1] = fmaxf(hi[1], q.y); hi[2] = fmaxf(hi[2], q.z);
    }
    for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = lo[k];
        s[3 + k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; k++) {
                s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + w]);
                s[3 + k][threadIdx.x] = fmaxf(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float mg = scenes[body_scene[b]].pair_margin;
        out[2 * b] = make_float4(s[0][0] - mg, s[1][0] - mg, s[2][0] - mg, 0.0f);
        out[2 * b + 1] = make_float4(s[3][0] + mg, s[4][0] + mg, s[5][0] + mg, 0.0f);
    }
}

#pragma clang fp contract(off)
__global__ void __launch_bounds__(256)
    k_bbody_from_fluid(const float4* __restrict__ rb_pos, const float4* __restrict__ rb_vel,
                       const float4* __restrict__ rb_cm, float4* __restrict__ out_f, float4* __restrict__ out_t,
                       uint32_t m, uint32_t nbodies, const uint32_t* __restrict__ body_scene,
                       const BatchBodyScene* __restrict__ scenes, const float4* __restrict__ sph_pos,
                       const float4* __restrict__ sph_vel, const uint32_t* __restrict__ sph_start,
                       const uint32_t* __restrict__ sph_end, uint32_t nsph) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = t / kBodyLanes;
    if (i >= m) return;  // the same for every lane of a group
    const int b = f2i_sat(rb_pos[i].w);
    if (b < 0 || (uint32_t)b >= nbodies) return;
    const BatchBodyScene& d = scenes[body_scene[b]];
    body_sphere_from_fluid<kBodyLanes>(i, rb_pos, rb_vel, rb_cm, out_f, out_t, nbodies, d.sphere_radius, sph_pos,
                                       sph_vel, sph_start, sph_end, nsph, d.fluid_radius, d.fluid, d.rb_to_fluid, d.ks,
                                       d.kd, t % kBodyLanes);
}

// Scenes with one body have no body-body pass (the single-fluid simulator
// does not launch it), so their spheres keep exactly the fluid's part.
__global__ void __launch_bounds__(256)
    k_bbody_from_bodies(const float4* __restrict__ rb_pos, const float4* __restrict__ rb_vel,
                        const float4* __restrict__ rb_cm, float4* __restrict__ out_f, float4* __restrict__ out_t,
                        uint32_t m, uint32_t nbodies, const uint32_t* __restrict__ body_scene,
                        const BatchBodyScene* __restrict__ scenes, const float4* __restrict__ body_bounds) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = t / kBodyLanes;
    if (i >= m) return;  // the same for every lane of a group
    const int b = f2i_sat(rb_pos[i].w);
    if (b < 0 || (uint32_t)b >= nbodies) return;
    const BatchBodyScene& d = scenes[body_scene[b]];
    if (d.body_end - d.body_begin < 2) return;
    BodyContact dyn = d.dyn;  // as fg_body_forces_from_bodies sets it up
    dyn.relation = make_int3(1, 1, 1);
    dyn.fluid_radius = 0.0f;
    dyn.eps = 0.001f;
    dyn.aabb = nullptr;
    const BodyContact st{};
    body_sphere_from_bodies<true, kBodyLanes>(i, rb_pos, rb_vel, rb_cm, out_f, out_t, nbodies, d.sphere_radius, dyn,
                                              st, body_bounds, d.body_begin, d.body_end, t % kBodyLanes);
}

__global__ void k_bintegrate_rb(float4* __restrict__ pos, const float* __restrict__ mass, float4* __restrict__ vel,
                                const float4* __restrict__ forces, uint32_t nb, const uint32_t* __restrict__ body_scene,
                                const BatchBodyScene* __restrict__ scenes, float4* __restrict__ L,
                                const float4* __restrict__ torque, const float4* __restrict__ inv_inertia,
                                float4* __restrict__ orient) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb) return;
    const BatchBodyScene& d = scenes[body_scene[idx]];
    integrate_body(idx, pos, mass, vel, forces, d.dt, d.ext, d.vmin, d.vmax, L, torque, inv_inertia, orient,
                   d.vlimit);
}
#pragma clang fp contract(fast)

int batch_body_transform(uint32_t* key, uint32_t* id, const float4* cm, const float4* orient, const float4* rel,
                         float4* world, uint32_t m, const uint32_t* body_scene, const BatchBodyScene* scenes) {
    if (m == 0) return FG_OK;
    hipLaunchKernelGGL(k_bbody_transform, dim3(grid_for(m, 256)), dim3(256), 0, state().stream, key, id, cm, orient,
                       rel, world, m, body_scene, scenes);
    return check_launch("batch_body_transform") ? FG_OK : fg_peek_error();
}

int batch_scene_bounds(const float4* world, const BatchBodyScene* scenes, uint32_t nscenes, float4* out) {
    if (nscenes == 0) return FG_OK;
    hipLaunchKernelGGL(k_bscene_bounds, dim3(nscenes), dim3(256), 0, state().stream, world, scenes, nscenes, out);
    return check_launch("batch_scene_bounds") ? FG_OK : fg_peek_error();
}

int batch_body_bounds(const float4* world, const uint32_t* offsets, const uint32_t* body_scene,
                      const BatchBodyScene* scenes, uint32_t nbodies, float4* out) {
    if (nbodies == 0) return FG_OK;
    hipLaunchKernelGGL(k_bbody_bounds, dim3(nbodies), dim3(256), 0, state().stream, world, offsets, body_scene,
                       scenes, nbodies, out);
    return check_launch("batch_body_bounds") ? FG_OK : fg_peek_error();
}

int batch_body_forces(const float4* world, const float4* vel, const float4* cm, float4* out_f, float4* out_t,
                      uint32_t m, uint32_t nbodies, const uint32_t* body_scene, const BatchBodyScene* scenes,
                      const float4* fluid_pos, const float4* fluid_vel, const uint32_t* fluid_start,
                      const uint32_t* fluid_end, uint32_t nfluid, const float4* body_bounds) {
    if (m == 0) return FG_OK;
    hipLaunchKernelGGL(k_bbody_from_fluid, dim3(grid_for((uint64_t)m * kBodyLanes, 256)), dim3(256), 0, state().stream, world, vel, cm,
                       out_f, out_t, m, nbodies, body_scene, scenes, fluid_pos, fluid_vel, fluid_start, fluid_end,
                       nfluid);
    hipLaunchKernelGGL(k_bbody_from_bodies, dim3(grid_for((uint64_t)m * kBodyLanes, 256)), dim3(256), 0, state().stream, world, vel, cm,
                       out_f, out_t, m, nbodies, body_scene, scenes, body_bounds);
    return check_launch("batch_body_forces") ? FG_OK : fg_peek_error();
}

int batch_integrate_bodies(float4* cm, const float* mass, float4* vel, const float4* force, uint32_t nbodies,
                           const uint32_t* body_scene, const BatchBodyScene* scenes, float4* L, const float4* torque,
                           const float4* inv_inertia, float4* orient) {
    if (nbodies == 0) return FG_OK;
    hipLaunchKernelGGL(k_bintegrate_rb, dim3(grid_for(nbodies, 64)), dim3(64), 0, state().stream, cm, mass, vel,
                       force, nbodies, body_scene, scenes, L, torque, inv_inertia, orient);
    return check_launch("batch_integrate_bodies") ? FG_OK : fg_peek_error();
}

static GridDesc body_grid(fg_float3 cs, uint32_t nb) {
    GridDesc g{};
    g.cs = make_float3(cs.x, cs.y, cs.z);
    g.nb = nb;
    g.mode = FG_KEY_REF_XOR;
    return g;
}

}  // namespace fg

using namespace fg;

// Contact description of one fg_body_grid (K_FluidSim.cu:372-386 arguments).
static BodyContact body_contact(const fg_body_grid& b) {
    BodyContact c{};
    c.enabled = b.numSpheres > 0;
    c.pos = reinterpret_cast<const float4*>(b.orderedSpheres);
    c.vel = reinterpret_cast<const float4*>(b.bodyVelocities);
    c.start = b.gridStart;
    c.end = b.gridEnd;
    c.aabb = reinterpret_cast<const float4*>(b.bounds);
    c.grid = body_grid(b.cellSize, b.numBuckets);
    c.relation = make_int3(b.relation.x, b.relation.y, b.relation.z);
    c.fluid_radius = b.fluidRadius;
    c.body_radius = b.sphereRadius;
    c.ks = b.kSpring;
    c.kd = b.kDamping;
    c.eps = 0.003f;
    c.count = b.numSpheres;
    c.nvel = b.numBodies;
    return c;
}

extern "C" {

// K_FluidSim.cu:615-680
void gSPHCalculateForcesWithSRB(fg_float4* oldPositions, fg_float4* oldVelocity, fg_float4* forces, fg_float4* newVel,
                                fg_float2* density_pressure, unsigned int* particleId, unsigned int* gridStart,
                                unsigned int* gridEnd, unsigned int numParticles, float mass, float kViscosity,
                                float _in_sph_particle_radius, fg_int3 _in_cell_size_relation,
                                float _in_srb_particle_radius, fg_float3 _in_srb_cell_size, float _in_kspring,
                                float _in_kdamping, fg_float4* _in_srb_positions, unsigned int* _in_srb_start_offset,
                                unsigned int* _in_srb_end_offset, unsigned int _in_srb_num_buckets,
                                fg_float4* _in_rb_positions, fg_float4* _in_rb_velocities,
                                unsigned int* _in_rb_start_offset, unsigned int* _in_rb_end_offset,
                                unsigned int _in_rb_num_buckets, bool useStaticRB, bool useDynamicRB) {
    (void)forces;
    if (numParticles == 0) return;
    if (!oldPositions || !oldVelocity || !newVel || !density_pressure || !particleId || !gridStart || !gridEnd) {
        set_error(FG_ERR_INVALID_ARGUMENT, "gSPHCalculateForcesWithSRB: null pointer");
        return;
    }
    if ((useStaticRB && (!_in_srb_positions || !_in_srb_start_offset || !_in_srb_end_offset ||
                         _in_srb_num_buckets == 0)) ||
        (useDynamicRB && (!_in_rb_positions || !_in_rb_velocities || !_in_rb_start_offset || !_in_rb_end_offset ||
                          _in_rb_num_buckets == 0))) {
        // The message echoes every by-value scalar of the call, so a caller
        // built against the reference prototypes can check what arrived.
        set_error(FG_ERR_INVALID_ARGUMENT,
                  "gSPHCalculateForcesWithSRB: body grid missing (numParticles=%u mass=%.9g kViscosity=%.9g "
                  "sph_particle_radius=%.9g cell_size_relation=(%d,%d,%d) srb_particle_radius=%.9g "
                  "srb_cell_size=(%.9g,%.9g,%.9g) kspring=%.9g kdamping=%.9g srb_num_buckets=%u rb_num_buckets=%u "
                  "useStaticRB=%d useDynamicRB=%d)",
                  numParticles, (double)mass, (double)kViscosity, (double)_in_sph_particle_radius,
                  _in_cell_size_relation.x, _in_cell_size_relation.y, _in_cell_size_relation.z,
                  (double)_in_srb_particle_radius, (double)_in_srb_cell_size.x, (double)_in_srb_cell_size.y,
                  (double)_in_srb_cell_size.z, (double)_in_kspring, (double)_in_kdamping, _in_srb_num_buckets,
                  _in_rb_num_buckets, (int)useStaticRB, (int)useDynamicRB);
        return;
    }
    Contacts c{};
    c.coupled = 1;
    const int3 rel = make_int3(_in_cell_size_relation.x, _in_cell_size_relation.y, _in_cell_size_relation.z);
    // The body-sphere arrays carry no length in this signature; the ranges in
    // the body grids bound every read.
    c.st.enabled = useStaticRB;
    c.st.pos = FG_F4(_in_srb_positions);
    c.st.vel = nullptr;
    c.st.start = _in_srb_start_offset;
    c.st.end = _in_srb_end_offset;
    c.st.grid = body_grid(_in_srb_cell_size, _in_srb_num_buckets);
    c.st.relation = rel;
    c.st.fluid_radius = _in_sph_particle_radius;
    c.st.body_radius = _in_srb_particle_radius;
    c.st.ks = _in_kspring;
    c.st.kd = _in_kdamping;
    c.st.eps = 0.003f;
    c.st.count = 0xFFFFFFFFu;
    c.st.nvel = 0;
    c.dyn = c.st;
    c.dyn.enabled = useDynamicRB;
    c.dyn.pos = FG_F4(_in_rb_positions);
    c.dyn.vel = FG_F4(_in_rb_velocities);
    c.dyn.start = _in_rb_start_offset;
    c.dyn.end = _in_rb_end_offset;
    c.dyn.grid = body_grid(_in_srb_cell_size, _in_rb_num_buckets);
    c.dyn.nvel = 0xFFFFFFFFu;
    launch_forces(FG_F4(oldPositions), FG_F4(oldVelocity), FG_F2(density_pressure), gridStart, gridEnd,
                  FG_F4(newVel), nullptr, particleId, numParticles, mass, kViscosity, nullptr, &c);
}

// Coupled forces with explicit sphere/body counts, optional neighbour masks
// (fg_sph_density_nbr), optional fused integration and optional sphere
// bounds used to skip the body walk far from every body.
int fg_sph_forces_coupled(fg_float4* orderedPositions, fg_float4* orderedVelocity, fg_float2* density_pressure,
                          unsigned int* gridStart, unsigned int* gridEnd, const unsigned long long* neighborMasks,
                          fg_float4* vel, fg_float4* pos, unsigned int* particleId, unsigned int numParticles,
                          float mass, float viscosity, const fg_body_grid* bodies,
                          const fg_body_grid* staticBodies) {
    if (numParticles == 0) return FG_OK;
    if (!orderedPositions || !orderedVelocity || !density_pressure || !gridStart || !gridEnd || !vel ||
        !particleId || (!bodies && !staticBodies)) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sph_forces_coupled: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (staticBodies && staticBodies->numSpheres > 0 &&
        (!staticBodies->orderedSpheres || !staticBodies->gridStart || !staticBodies->gridEnd ||
         staticBodies->numBuckets == 0)) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sph_forces_coupled: static body grid incomplete");
        return FG_ERR_INVALID_ARGUMENT;
    }
    Contacts c{};
    c.coupled = 1;
    if (staticBodies) {
        c.st = body_contact(*staticBodies);
        c.st.vel = nullptr;  // static spheres do not move
        c.st.nvel = 0;
    }
    if (bodies) c.dyn = body_contact(*bodies);
    return launch_forces(FG_F4(orderedPositions), FG_F4(orderedVelocity), FG_F2(density_pressure), gridStart, gridEnd,
                         FG_F4(vel), pos ? FG_F4(pos) : nullptr, particleId, numParticles, mass, viscosity,
                         neighborMasks, &c);
}

// K_RigidBodySim.cu:309-342
void gCalculateRigidBodyForcesFromFluids(fg_float4* _in_RBPositions, fg_float4* _in_RBVelocites,
                                         fg_float4* _in_RBCMPos, fg_float4* _out_Forces, fg_float4* _out_torque,
                                         unsigned int* _in_RBGridStart, unsigned int* _in_RBGridEnd,
                                         unsigned int _in_NumParticles, fg_float3 _in_CellSize, float _in_Radius,
                                         int _in_RBHashTableSize, fg_float4* _in_sph_particles,
                                         fg_float4* _in_sph_velocities, unsigned int* _in_sph_grid_start,
                                         unsigned int* _in_sph_grid_end, float _in_sph_particle_radius,
                                         fg_float3 _in_sph_cell_size, unsigned int _in_sph_num_buckets,
                                         fg_int3 _in_cell_size_relation, float _in_sph_kSpring,
                                         float _in_sph_kDamping) {
    (void)_in_RBGridStart;
    (void)_in_RBGridEnd;
    (void)_in_CellSize;
    (void)_in_RBHashTableSize;
    fg_body_forces_from_fluid(_in_RBPositions, _in_RBVelocites, _in_RBCMPos, _out_Forces, _out_torque,
                              _in_NumParticles, 0xFFFFFFFFu, _in_Radius, _in_sph_particles, _in_sph_velocities,
                              _in_sph_grid_start, _in_sph_grid_end, 0xFFFFFFFFu, _in_sph_particle_radius,
                              _in_sph_cell_size, _in_sph_num_buckets, _in_cell_size_relation, _in_sph_kSpring,
                              _in_sph_kDamping);
}

int fg_body_forces_from_fluid(fg_float4* spheres, fg_float4* bodyVelocities, fg_float4* bodyCM, fg_float4* outForces,
                              fg_float4* outTorques, unsigned int numSpheres, unsigned int numBodies, float radius,
                              fg_float4* fluidOrderedPositions, fg_float4* fluidOrderedVelocities,
                              unsigned int* fluidGridStart, unsigned int* fluidGridEnd, unsigned int numFluid,
                              float fluidRadius, fg_float3 fluidCellSize, unsigned int fluidNumBuckets,
                              fg_int3 relation, float kSpring, float kDamping) {
    if (numSpheres == 0) return FG_OK;
    if (!spheres || !bodyVelocities || !bodyCM || !outForces || !outTorques || !fluidOrderedPositions ||
        !fluidOrderedVelocities || !fluidGridStart || !fluidGridEnd) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_body_forces_from_fluid: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    const State& st = state();
    // The fluid grid is walked with the key function it was built with.
    const GridDesc g = fluid_grid_desc(st.common, fluidCellSize, fluidNumBuckets, st.key_mode);
    hipLaunchKernelGGL(k_body_forces_from_fluid, dim3(grid_for((uint64_t)numSpheres * kBodyLanes, 256)), dim3(256), 0, st.stream,
                       FG_F4(spheres), FG_F4(bodyVelocities), FG_F4(bodyCM), FG_F4(outForces), FG_F4(outTorques),
                       numSpheres, numBodies, radius, FG_F4(fluidOrderedPositions), FG_F4(fluidOrderedVelocities),
                       fluidGridStart, fluidGridEnd, numFluid, fluidRadius, g,
                       make_int3(relation.x, relation.y, relation.z), kSpring, kDamping);
    return check_launch("fg_body_forces_from_fluid") ? FG_OK : fg_peek_error();
}

// K_RigidBodySim.cu:280-305.  The sphere array is both the query set and the
// (sorted) grid array, as in the reference; the body count is unknown, so the
// velocity and centre reads are bounded by the caller's data only.
void gCalculateRigidBodyForcesFromRB(fg_float4* _in_RBPositions, fg_float4* _in_RBVelocites,
                                     fg_float4* _in_RBCMPosition, fg_float4* _out_torque, fg_float4* _out_Forces,
                                     unsigned int* _in_RBGridStart, unsigned int* _in_RBGridEnd,
                                     unsigned int _in_NumParticles, fg_float3 _in_CellSize, float _in_Radius,
                                     int _in_RBHashTableSize, int _in_SRBHashTableSize, fg_float4* _in_SRBPositions,
                                     unsigned int* _in_SRBGridStart, unsigned int* _in_SRBGridEnd,
                                     float _in_RBkSpring, float _in_RBkDamping, int _in_numSRB) {
    if (_in_NumParticles == 0) return;
    if (!_in_RBPositions || !_in_RBVelocites || !_in_RBCMPosition || !_out_torque || !_out_Forces ||
        !_in_RBGridStart || !_in_RBGridEnd || _in_RBHashTableSize <= 0 ||
        (_in_numSRB > 0 && (!_in_SRBPositions || !_in_SRBGridStart || !_in_SRBGridEnd || _in_SRBHashTableSize <= 0))) {
        set_error(FG_ERR_INVALID_ARGUMENT, "gCalculateRigidBodyForcesFromRB: null pointer or empty grid");
        return;
    }
    BodyContact dyn{};
    dyn.enabled = 1;
    dyn.pos = FG_F4(_in_RBPositions);
    dyn.vel = FG_F4(_in_RBVelocites);
    dyn.start = _in_RBGridStart;
    dyn.end = _in_RBGridEnd;
    dyn.grid = body_grid(_in_CellSize, (uint32_t)_in_RBHashTableSize);
    dyn.relation = make_int3(1, 1, 1);
    dyn.body_radius = _in_Radius;
    dyn.ks = _in_RBkSpring;
    dyn.kd = _in_RBkDamping;
    dyn.eps = 0.001f;
    dyn.count = _in_NumParticles;
    dyn.nvel = 0xFFFFFFFFu;
    BodyContact st = dyn;
    st.enabled = _in_numSRB > 0;
    st.pos = FG_F4(_in_SRBPositions);
    st.vel = nullptr;
    st.start = _in_SRBGridStart;
    st.end = _in_SRBGridEnd;
    st.grid = body_grid(_in_CellSize, _in_numSRB > 0 ? (uint32_t)_in_SRBHashTableSize : 1u);
    st.count = 0xFFFFFFFFu;
    st.nvel = 0;
    hipLaunchKernelGGL(k_body_forces_from_bodies<false>, dim3(grid_for((uint64_t)_in_NumParticles * kBodyLanes, 256)), dim3(256), 0,
                       state().stream, FG_F4(_in_RBPositions), FG_F4(_in_RBVelocites), FG_F4(_in_RBCMPosition),
                       FG_F4(_out_Forces), FG_F4(_out_torque), _in_NumParticles, 0xFFFFFFFFu, _in_Radius, dyn, st,
                       (const float4*)nullptr, 0u);
    check_launch("gCalculateRigidBodyForcesFromRB");
}

int fg_body_forces_from_bodies(const fg_float4* spheres, const fg_float4* bodyVelocities, const fg_float4* bodyCM,
                               fg_float4* outForces, fg_float4* outTorques, unsigned int numSpheres,
                               unsigned int numBodies, float radius, const fg_body_grid* bodies,
                               const fg_float4* bodyBounds, const fg_body_grid* staticBodies, int accumulate) {
    if (numSpheres == 0) return FG_OK;
    if (!spheres || !bodyVelocities || !bodyCM || !outForces || !outTorques || (!bodies && !staticBodies)) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_body_forces_from_bodies: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    for (const fg_body_grid* g : {bodies, staticBodies})
        if (g && g->numSpheres > 0 && (!g->orderedSpheres || !g->gridStart || !g->gridEnd || g->numBuckets == 0)) {
            set_error(FG_ERR_INVALID_ARGUMENT, "fg_body_forces_from_bodies: body grid incomplete");
            return FG_ERR_INVALID_ARGUMENT;
        }
    if (bodies && bodies->numSpheres > 0 && !bodies->bodyVelocities) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_body_forces_from_bodies: dynamic grid without body velocities");
        return FG_ERR_INVALID_ARGUMENT;
    }
    BodyContact dyn{}, st{};
    if (bodies) {
        dyn = body_contact(*bodies);
        dyn.eps = 0.001f;
        dyn.aabb = nullptr;
    }
    if (staticBodies) {
        st = body_contact(*staticBodies);
        st.eps = 0.001f;
        st.vel = nullptr;
        st.nvel = 0;
    }
    const float4* bb = bodies ? reinterpret_cast<const float4*>(bodyBounds) : nullptr;
    if (accumulate)
        hipLaunchKernelGGL(k_body_forces_from_bodies<true>, dim3(grid_for((uint64_t)numSpheres * kBodyLanes, 256)), dim3(256), 0,
                           state().stream, reinterpret_cast<const float4*>(spheres),
                           reinterpret_cast<const float4*>(bodyVelocities), reinterpret_cast<const float4*>(bodyCM),
                           FG_F4(outForces), FG_F4(outTorques), numSpheres, numBodies, radius, dyn, st, bb,
                           numBodies);
    else
        hipLaunchKernelGGL(k_body_forces_from_bodies<false>, dim3(grid_for((uint64_t)numSpheres * kBodyLanes, 256)), dim3(256), 0,
                           state().stream, reinterpret_cast<const float4*>(spheres),
                           reinterpret_cast<const float4*>(bodyVelocities), reinterpret_cast<const float4*>(bodyCM),
                           FG_F4(outForces), FG_F4(outTorques), numSpheres, numBodies, radius, dyn, st, bb,
                           numBodies);
    return check_launch("fg_body_forces_from_bodies") ? FG_OK : fg_peek_error();
}

int fg_body_bounds(const fg_float4* spheres, const unsigned int* bodyOffsets, unsigned int numBodies, float margin,
                   fg_float4* out2) {
    if (numBodies == 0) return FG_OK;
    if (!spheres || !bodyOffsets || !out2) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_body_bounds: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    hipLaunchKernelGGL(k_body_bounds, dim3(numBodies), dim3(256), 0, state().stream,
                       reinterpret_cast<const float4*>(spheres), bodyOffsets, numBodies, margin, FG_F4(out2));
    return check_launch("fg_body_bounds") ? FG_OK : fg_peek_error();
}

// K_RigidBodySim.cu:344-352
void gCopyRBDataToParticle(unsigned int* _out_force_rb, unsigned int* _out_force_index, fg_float4* _in_rb_particle_forces,
                           size_t num_particles) {
    if (num_particles == 0) return;
    hipLaunchKernelGGL(k_copy_rb_data, dim3(grid_for(num_particles, 256)), dim3(256), 0, state().stream,
                       _out_force_rb, _out_force_index, FG_F4(_in_rb_particle_forces), (uint64_t)num_particles);
    check_launch("gCopyRBDataToParticle");
}

// K_RigidBodySim.cu:354-366
void gReorderForceTorque(fg_float4* _out_force, fg_float4* _out_torque, fg_float4* _in_force, fg_float4* _in_torque,
                         unsigned int* _in_ordered_index, unsigned int _num_values) {
    if (_num_values == 0) return;
    hipLaunchKernelGGL(k_reorder_force_torque, dim3(grid_for(_num_values, 256)), dim3(256), 0, state().stream,
                       FG_F4(_out_force), FG_F4(_out_torque), FG_F4(_in_force), FG_F4(_in_torque), _in_ordered_index,
                       _num_values);
    check_launch("gReorderForceTorque");
}

// K_RigidBodySim.cu:368-377
void gAddForceTorque(fg_float4* _in_sum_force, fg_float4* _in_sum_torque, fg_float4* _in_ordened_force,
                     fg_float4* _out_to_force, fg_float4* _out_to_torque, int* _in_sum_offset, size_t num_rigid_body) {
    if (num_rigid_body == 0) return;
    hipLaunchKernelGGL(k_add_force_torque, dim3(grid_for(num_rigid_body, 256)), dim3(256), 0, state().stream,
                       FG_F4(_in_sum_force), FG_F4(_in_sum_torque), FG_F4(_in_ordened_force), FG_F4(_out_to_force),
                       FG_F4(_out_to_torque), _in_sum_offset, (uint64_t)num_rigid_body);
    check_launch("gAddForceTorque");
}

// K_RigidBodySim.cu:381-392
void gIntegrateRB(fg_float4* _in_RBPosition, float* _in_mass, fg_float4* _in_RBVelocities, fg_float4* _in_rbforces,
                  size_t num_rb, float _elapsed_time_in_seconds, fg_float4 _in_external_force, fg_float3 min_volume,
                  fg_float3 max_volume, fg_float4* _angular_momentum, fg_float4* _torque, fg_float4* _inertia_tensor,
                  fg_float4* _orientation, fg_float4* _inv_inertia_tensor, float velocity_limit) {
    (void)_inertia_tensor;
    if (num_rb == 0) return;
    hipLaunchKernelGGL(k_integrate_rb, dim3(grid_for(num_rb, 64)), dim3(64), 0, state().stream, FG_F4(_in_RBPosition),
                       _in_mass, FG_F4(_in_RBVelocities), FG_F4(_in_rbforces), (uint64_t)num_rb,
                       _elapsed_time_in_seconds,
                       make_float4(_in_external_force.x, _in_external_force.y, _in_external_force.z,
                                   _in_external_force.w),
                       make_float3(min_volume.x, min_volume.y, min_volume.z),
                       make_float3(max_volume.x, max_volume.y, max_volume.z), FG_F4(_angular_momentum),
                       FG_F4(_torque), FG_F4(_inv_inertia_tensor), FG_F4(_orientation), velocity_limit);
    check_launch("gIntegrateRB");
}

int fg_body_reduce_f64(const fg_float4* forces, const fg_float4* torques, const unsigned int* bodyOffsets,
                       unsigned int numBodies, fg_double4* outForces, fg_double4* outTorques, fg_float4* accForces,
                       fg_float4* accTorques) {
    if (numBodies == 0) return FG_OK;
    if (!forces || !torques || !bodyOffsets || !outForces || !outTorques) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_body_reduce_f64: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    hipLaunchKernelGGL(k_body_reduce_f64, dim3(numBodies), dim3(kBodySumThreads), 0, state().stream,
                       reinterpret_cast<const float4*>(forces), reinterpret_cast<const float4*>(torques), bodyOffsets,
                       numBodies, reinterpret_cast<double4*>(outForces), reinterpret_cast<double4*>(outTorques),
                       accForces ? FG_F4(accForces) : nullptr, accTorques ? FG_F4(accTorques) : nullptr);
    return check_launch("fg_body_reduce_f64") ? FG_OK : fg_peek_error();
}

int fg_body_mass_properties_f64(const fg_float4* worldSpheres, const unsigned int* bodyOffsets, unsigned int numBodies,
                                const double* masses, double* centreOfMass, double* inertia) {
    if (numBodies == 0) return FG_OK;
    if (!worldSpheres || !bodyOffsets || !masses || !centreOfMass || !inertia) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_body_mass_properties_f64: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    for (int pass = 0; pass < 2; pass++)
        hipLaunchKernelGGL(k_body_mass_props_f64, dim3(numBodies), dim3(kReduceThreads), 0, state().stream,
                           reinterpret_cast<const float4*>(worldSpheres), bodyOffsets, numBodies, masses,
                           centreOfMass, inertia, pass);
    return check_launch("fg_body_mass_properties_f64") ? FG_OK : fg_peek_error();
}

int fg_sphere_bounds(const fg_float4* spheres, unsigned int numSpheres, float margin, fg_float4* out2) {
    if (!spheres || !out2) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sphere_bounds: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    hipLaunchKernelGGL(k_sphere_bounds, dim3(1), dim3(1024), 0, state().stream, reinterpret_cast<const float4*>(spheres),
                       numSpheres, margin, FG_F4(out2));
    return check_launch("fg_sphere_bounds") ? FG_OK : fg_peek_error();
}

}  // extern "C"
