/*
 * fg_oracle.cpp -- CPU restatement of FluidGrid's per-timestep SPH step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * fluidgrid_amd/: tests, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it; the product path never does.
 *
 * Every function restates one reference routine (paths relative to the
 * FluidGrid source tree) and is written from the reference's arithmetic as
 * read in the source, with -ffp-contract=off so that every multiply and add
 * rounds separately as the source spells it:
 *
 *   hash (REF_XOR)         GPU_Kernels/K_DeviceFunctions.cu:17-26
 *   hash kernel            GPU_Kernels/K_HashGrid.cu:26-38
 *   body-sphere hash       GPU_Kernels/K_HashGrid.cu:43-63, K_MathUtils.cu:57-181
 *   sort                   GPU_Kernels/K_FluidSim.cu:685-691 (stable key/value)
 *   cell offsets, reorder  GPU_Kernels/K_HashGrid.cu:70-133, 140-197, 254-255
 *   density                GPU_Kernels/K_FluidSim.cu:46-79, 210-254
 *   pair forces            GPU_Kernels/K_FluidSim.cu:85-205
 *   forces kernel          GPU_Kernels/K_FluidSim.cu:260-309
 *   walls                  GPU_Kernels/K_Boundaries_Walls.cu:8-125,
 *                          GPU_Kernels/K_Boundaries_Common.cu:12-82
 *   integrate              GPU_Kernels/K_FluidSim.cu:418-438
 *   penalty contact        GPU_Kernels/K_DeviceFunctions.cu:32-159
 *   coupled forces         GPU_Kernels/K_FluidSim.cu:315-413
 *   body forces from fluid GPU_Kernels/K_RigidBodySim.cu:66-99
 *   body reduction         GPU_Kernels/K_RigidBodySim.cu:102-155,
 *                          Simulators/RigidBodySimulator.cpp:522-566
 *   body integration       GPU_Kernels/K_RigidBodySim.cu:160-262,
 *                          GPU_Kernels/K_MathUtils.cu:160-236
 *   kernel constants       HeterogeneousSimulator.cpp:33-39
 *
 * The LINEAR key mode is this build's own (a linear cell index with an
 * integer +-1 stencil); it is restated here with the same formula the HIP
 * kernels use.
 *
 * Density and forces are templated on the arithmetic type: Real=float is
 * the FP32 restatement, Real=double the FP64 oracle step (same inputs, same
 * parameter values promoted to double, every operation in double).
 *
 * Pinning: tests/test_oracle_*.py check each function against closed forms,
 * invariants and O(N^2) brute force (the reference ships no tests or golden
 * vectors).  Functions without such a check are named "parity unpinned" in
 * the test module headers.
 */
#include "../include/fluidgrid_amd.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

static_assert(sizeof(fg_common_fluid_params) == 104, "CommonFluidParams layout");
static_assert(sizeof(fg_fluid_params) == 28, "FluidParams layout");
static_assert(sizeof(fg_kernels_cache) == 12, "KernelsCache layout");
static_assert(sizeof(fg_common_rigid_body_params) == 112, "CommonRigidBodyParams layout");

namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;

/* When set, every bucket is walked from its last particle to its first: an
 * equally valid FP32 evaluation whose divergence from the forward order
 * calibrates trajectory tolerances. */
int g_reverse = 0;

inline uint32_t walk(uint32_t s, uint32_t e, uint32_t jj) { return g_reverse ? (e - 1 - (jj - s)) : jj; }

/* float -> int conversion with the GPU's semantics (truncate, saturate,
 * NaN -> 0); plain C++ conversion is undefined outside the int range. */
inline int32_t f2i_sat(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}

/* K_DeviceFunctions.cu:17-26 */
inline uint32_t hash_ref_xor(float px, float py, float pz, fg_float3 cs, uint32_t tableSize) {
    int32_t x = f2i_sat((px + 5000.0f) / cs.x);
    int32_t y = f2i_sat((py + 5000.0f) / cs.y);
    int32_t z = f2i_sat((pz + 5000.0f) / cs.z);
    uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
    return h % tableSize;
}

inline int cell_coord(float p, float vmin, float cs, int n) {
    int32_t c = f2i_sat(floorf((p - vmin) / cs));
    return c < 0 ? 0 : (c > n - 1 ? n - 1 : c);
}

struct Cell { int x, y, z; };

inline Cell linear_cell(const fg_float4& p, const fg_common_fluid_params& cp) {
    return {cell_coord(p.x, cp.volume_min.x, cp.cell_size.x, cp.bucket_size.x),
            cell_coord(p.y, cp.volume_min.y, cp.cell_size.y, cp.bucket_size.y),
            cell_coord(p.z, cp.volume_min.z, cp.cell_size.z, cp.bucket_size.z)};
}

inline uint32_t linear_key(const Cell& c, const fg_common_fluid_params& cp) {
    return (uint32_t)c.x + (uint32_t)cp.bucket_size.x *
           ((uint32_t)c.y + (uint32_t)cp.bucket_size.y * (uint32_t)c.z);
}

inline uint32_t num_buckets(const fg_common_fluid_params& cp) {
    return (uint32_t)cp.bucket_size.x * (uint32_t)cp.bucket_size.y * (uint32_t)cp.bucket_size.z;
}

/* The 27 stencil buckets of a particle in z, y, x loop order
 * (K_FluidSim.cu:231-246).  REF_XOR rehashes p + k*cs and may repeat a
 * bucket; LINEAR skips cells outside the grid (count < 27). */
inline int stencil(const fg_float4& p, const fg_common_fluid_params& cp, int mode, uint32_t out[27]) {
    int m = 0;
    if (mode == FG_KEY_REF_XOR) {
        const uint32_t nb = num_buckets(cp);
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++) {
                    fg_float3 off = cp.cell_size;
                    off.x *= (float)x; off.y *= (float)y; off.z *= (float)z;
                    out[m++] = hash_ref_xor(p.x + off.x, p.y + off.y, p.z + off.z, cp.cell_size, nb);
                }
    } else {
        Cell c = linear_cell(p, cp);
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++) {
                    Cell d{c.x + x, c.y + y, c.z + z};
                    if (d.x < 0 || d.y < 0 || d.z < 0 || d.x >= cp.bucket_size.x ||
                        d.y >= cp.bucket_size.y || d.z >= cp.bucket_size.z)
                        continue;
                    out[m++] = linear_key(d, cp);
                }
    }
    return m;
}

template <class R> struct V3 { R x, y, z; };
template <class R> inline V3<R> v3(R x, R y, R z) { return {x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline V3<R>& operator+=(V3<R>& a, V3<R> b) { a = a + b; return a; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R len(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> inline V3<R> xyz(const fg_float4& p) { return {(R)p.x, (R)p.y, (R)p.z}; }

template <class R> inline R rpow3(R x) { return std::pow(x, (R)3); }

/* K_Boundaries_Common.cu:12-23 */
template <class R>
inline V3<R> repulsion(V3<R> vel, V3<R> n, R dist, R damping, R stiffness) {
    return (stiffness * dist - damping * dot(n, vel)) * n;
}

/* K_Boundaries_Common.cu:48-82 (f_n, f_t are computed there but unused) */
template <class R>
inline V3<R> friction(V3<R> vel, V3<R> /*force*/, V3<R> n, R kinetic, R static_limit) {
    V3<R> v_n = vel * dot(n, vel);
    V3<R> v_t = vel - v_n;
    V3<R> neg = v3<R>(-v_t.x, -v_t.y, -v_t.z);
    if ((v_t.x + v_t.y + v_t.z) / (R)3 > static_limit) return neg;
    return kinetic * neg;
}

/* K_Boundaries_Walls.cu:8-65; faces y-, y+, z-, z+, x-, x+ */
template <class R>
V3<R> walls_no_penetration(V3<R> pos, V3<R> vel, V3<R> gmin, V3<R> gmax, R bd, R stiff, R damp, R s) {
    const R eps = (R)0.00001f;
    V3<R> f = v3<R>(0, 0, 0);
    R d;
    d = bd - ((pos.y - gmin.y) * s);
    if (d > eps) f += repulsion(vel, v3<R>(0, 1, 0), d, damp, stiff);
    d = bd - ((gmax.y - pos.y) * s);
    if (d > eps) f += repulsion(vel, v3<R>(0, -1, 0), d, damp, stiff);
    d = bd - ((pos.z - gmin.z) * s);
    if (d > eps) f += repulsion(vel, v3<R>(0, 0, 1), d, damp, stiff);
    d = bd - ((gmax.z - pos.z) * s);
    if (d > eps) {
        V3<R> n = v3<R>(0, 0, -1);
        R adj = stiff * d - damp * dot(n, vel);
        f += adj * n;
    }
    d = bd - ((pos.x - gmin.x) * s);
    if (d > eps) f += repulsion(vel, v3<R>(1, 0, 0), d, damp, stiff);
    d = bd - ((gmax.x - pos.x) * s);
    if (d > eps) f += repulsion(vel, v3<R>(-1, 0, 0), d, damp, stiff);
    return f;
}

/* K_Boundaries_Walls.cu:68-125 */
template <class R>
V3<R> walls_no_slip(V3<R> pos, V3<R> vel, V3<R> force, V3<R> gmin, V3<R> gmax, R bd, R kin, R stat, R s) {
    const R eps = (R)0.00001f;
    V3<R> f = v3<R>(0, 0, 0);
    R d;
    d = bd - ((pos.y - gmin.y) * s);
    if (d > eps) f += friction(vel, force, v3<R>(0, 1, 0), kin, stat);
    d = bd - ((gmax.y - pos.y) * s);
    if (d > eps) f += friction(vel, force, v3<R>(0, -1, 0), kin, stat);
    d = bd - ((pos.z - gmin.z) * s);
    if (d > eps) f += friction(vel, force, v3<R>(0, 0, 1), kin, stat);
    d = bd - ((gmax.z - pos.z) * s);
    if (d > eps) f += friction(vel, force, v3<R>(0, 0, -1), kin, stat);
    d = bd - ((pos.x - gmin.x) * s);
    if (d > eps) f += friction(vel, force, v3<R>(1, 0, 0), kin, stat);
    d = bd - ((gmax.x - pos.x) * s);
    if (d > eps) f += friction(vel, force, v3<R>(-1, 0, 0), kin, stat);
    return f;
}

/* The wall terms of K_FluidSim.cu:289-295 applied to a total force. */
template <class R>
V3<R> wall_terms(const fg_float4& pos4, const fg_float4& vel4, V3<R> total, const fg_common_fluid_params& cp) {
    V3<R> pos = xyz<R>(pos4), vel = xyz<R>(vel4);
    const R s = (R)cp.simulation_scale;
    const R bd = (R)0.003043174f;
    V3<R> w = walls_no_penetration<R>(pos, vel, v3<R>(cp.volume_min.x, cp.volume_min.y, cp.volume_min.z),
                                      v3<R>(cp.volume_max.x, cp.volume_max.y, cp.volume_max.z), bd,
                                      (R)20000, (R)256, s);
    w += walls_no_slip<R>(pos, vel, total + w, v3<R>(-11.0f, -40.0f, -11.0f), v3<R>(11.0f, 200.0f, 11.0f),
                          bd, (R)0, (R)0, s);
    return w;
}

} // namespace

extern "C" {

/* ---- kernel constants (HeterogeneousSimulator.cpp:33-39) ---------------- */
void orc_kernels_cache(float h, fg_kernels_cache* out) {
    const float PI = 3.14159265f;
    out->kPoly6 = (float)(315.0f / (64.0f * PI * std::pow((double)h, 9)));
    out->kSpikyGradient = (float)(-(45.0f / (PI * std::pow((double)h, 6))));
    out->kViscosityLaplacian = (float)(45.0f / (PI * std::pow((double)h, 6)));
}

/* ---- hash (K_HashGrid.cu:26-38) ---------------------------------------- */
void orc_hash(const fg_float4* pos, uint32_t n, uint32_t numBuckets, fg_float3 cs,
              const fg_common_fluid_params* cp, int mode, uint32_t* key, uint32_t* id) {
    for (uint32_t i = 0; i < n; i++) {
        id[i] = i;
        if (mode == FG_KEY_REF_XOR)
            key[i] = hash_ref_xor(pos[i].x, pos[i].y, pos[i].z, cs, numBuckets);
        else
            key[i] = linear_key(linear_cell(pos[i], *cp), *cp);
    }
}

/* ---- stable key/value sort (K_FluidSim.cu:685-691) ---------------------
 * A counting sort placing elements in index order is stable by
 * construction; keys >= 2^26 fall back to std::stable_sort. */
void orc_sort_pairs(uint32_t* key, uint32_t* val, uint32_t n) {
    if (n == 0) return;
    uint32_t kmax = *std::max_element(key, key + n);
    std::vector<uint32_t> k2(n), v2(n);
    if (kmax < (1u << 26)) {
        std::vector<uint32_t> off((size_t)kmax + 2, 0);
        for (uint32_t i = 0; i < n; i++) off[(size_t)key[i] + 1]++;
        for (size_t b = 1; b < off.size(); b++) off[b] += off[b - 1];
        for (uint32_t i = 0; i < n; i++) {
            uint32_t d = off[key[i]]++;
            k2[d] = key[i];
            v2[d] = val[i];
        }
    } else {
        std::vector<uint32_t> idx(n);
        std::iota(idx.begin(), idx.end(), 0u);
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
        for (uint32_t i = 0; i < n; i++) { k2[i] = key[idx[i]]; v2[i] = val[idx[i]]; }
    }
    std::memcpy(key, k2.data(), sizeof(uint32_t) * n);
    std::memcpy(val, v2.data(), sizeof(uint32_t) * n);
}

/* ---- cell offsets + reorder (K_HashGrid.cu:70-133, 254-255) ------------ */
void orc_cell_offsets(uint32_t* start, uint32_t* end, const uint32_t* key, const uint32_t* id,
                      const fg_float4* pos, fg_float4* sortedPos, const fg_float4* vel,
                      fg_float4* sortedVel, uint32_t n, uint32_t numBuckets) {
    std::fill(start, start + numBuckets, kEmpty);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t h = key[i];
        if (i == 0 || h != key[i - 1]) {
            start[h] = i;
            if (i > 0) end[key[i - 1]] = i;
        }
        if (i == n - 1) end[h] = i + 1;
        sortedPos[i] = pos[id[i]];
        if (vel) sortedVel[i] = vel[id[i]];
    }
}

} // extern "C"

/* ---- density (K_FluidSim.cu:46-79, 210-254) ---------------------------
 * out2 receives (max(1,rho), k*(rho-rho0)) for w > 0; other entries are
 * left untouched.  Real=double accumulates in double and writes doubles. */
template <class R, class Out>
static void density_t(const fg_float4* sp, Out* rho_p, const uint32_t* start, const uint32_t* end,
                      uint32_t n, float k, float rho0, float m, const fg_common_fluid_params& cp,
                      const fg_kernels_cache& kc, int mode) {
    const R scale = (R)cp.simulation_scale, h = (R)cp.kernel_radius, h2 = (R)cp.kernel_radius_squared;
    const R mkp = (R)m * (R)kc.kPoly6;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const uint32_t i = (uint32_t)ii;
        const fg_float4 p = sp[i];
        if (!(p.w > 0.0f)) continue;
        uint32_t b[27];
        int nb = stencil(p, cp, mode, b);
        R rho = 0;
        for (int c = 0; c < nb; c++) {
            R part = 0;
            uint32_t s = start[b[c]];
            if (s != kEmpty) {
                V3<R> p1 = xyz<R>(p) * scale;
                uint32_t e = end[b[c]];
                for (uint32_t jj = s; jj < e; jj++) {
                    const uint32_t j = walk(s, e, jj);
                    V3<R> p2 = xyz<R>(sp[j]) * scale;
                    V3<R> d = p1 - p2;
                    R l = len(d);
                    if (l < h) part += mkp * rpow3<R>(h2 - (l * l));
                }
            }
            rho += part;
        }
        rho_p[2 * i + 0] = (Out)std::max((R)1, rho);
        rho_p[2 * i + 1] = (Out)((R)k * (rho - (R)rho0));
    }
}

extern "C" {

void orc_density_f32(const fg_float4* sp, float* rho_p, const uint32_t* start, const uint32_t* end,
                     uint32_t n, float k, float rho0, float m, const fg_common_fluid_params* cp,
                     const fg_kernels_cache* kc, int mode) {
    density_t<float, float>(sp, rho_p, start, end, n, k, rho0, m, *cp, *kc, mode);
}

void orc_density_f64(const fg_float4* sp, double* rho_p, const uint32_t* start, const uint32_t* end,
                     uint32_t n, float k, float rho0, float m, const fg_common_fluid_params* cp,
                     const fg_kernels_cache* kc, int mode) {
    density_t<double, double>(sp, rho_p, start, end, n, k, rho0, m, *cp, *kc, mode);
}

} // extern "C"

/* ---- forces (K_FluidSim.cu:85-205, 260-309) ----------------------------
 * Writes v_out[id[i]] = sortedVel[i] + a_i*dt for every i (a = 0 when
 * w <= 0).  acc_out (optional, sorted order, xyz) receives a_i after walls
 * and clamp; mag_out (optional) receives m*sum_j(|pressure_ij|+|visc_ij|),
 * the magnitude scale of the pair sum used by the parity tolerance.
 * rho_p holds (rho, p) pairs of type In. */
template <class R, class In>
static void forces_t(const fg_float4* sp, const fg_float4* sv, const In* rho_p, const uint32_t* start,
                     const uint32_t* end, R* v_out, const uint32_t* id, uint32_t n, float m, float mu,
                     const fg_common_fluid_params& cp, const fg_kernels_cache& kc, int mode,
                     R* acc_out, R* mag_out) {
    const R scale = (R)cp.simulation_scale, h = (R)cp.kernel_radius;
    const R kSpiky = (R)kc.kSpikyGradient, kVisc = (R)kc.kViscosityLaplacian;
    const R vlim = (R)cp.velocity_limit, dt = (R)cp.deltaTimeIntegration;
    const V3<R> g = v3<R>(cp.global_external_force.x, cp.global_external_force.y, cp.global_external_force.z);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const uint32_t i = (uint32_t)ii;
        const fg_float4 pos = sp[i], vel = sv[i];
        V3<R> total = v3<R>(0, 0, 0);
        R mag = 0;
        if (pos.w > 0.0f) {
            uint32_t b[27];
            int nb = stencil(pos, cp, mode, b);
            const R r1 = (R)rho_p[2 * i], pr1 = (R)rho_p[2 * i + 1];
            const V3<R> v1 = xyz<R>(vel);
            V3<R> internal = v3<R>(0, 0, 0);
            for (int c = 0; c < nb; c++) {
                V3<R> part = v3<R>(0, 0, 0);
                uint32_t s = start[b[c]];
                if (s != kEmpty) {
                    V3<R> p1 = xyz<R>(pos) * scale;
                    uint32_t e = end[b[c]];
                    for (uint32_t jj = s; jj < e; jj++) {
                        const uint32_t j = walk(s, e, jj);
                        if (j == i) continue;
                        V3<R> p2 = xyz<R>(sp[j]) * scale;
                        const R r2 = (R)rho_p[2 * j], pr2 = (R)rho_p[2 * j + 1];
                        const V3<R> v2 = xyz<R>(sv[j]);
                        /* SPHCalculatePressureForceD, K_FluidSim.cu:107-120 */
                        V3<R> fp = v3<R>(0, 0, 0);
                        {
                            R pp = (pr1 + pr2) / (r1 * r2);
                            V3<R> diff = p1 - p2;
                            R l = len(diff);
                            if (!(l > h)) {
                                R r = kSpiky * ((h - l) * (h - l));
                                fp = (diff / l) * r * pp * (R)-1.0f;
                            }
                        }
                        part += fp;
                        /* SPHCalculateViscosityForceD, K_FluidSim.cu:85-102 */
                        V3<R> fv = v3<R>(0, 0, 0);
                        {
                            V3<R> diff = p1 - p2;
                            R l = len(diff);
                            if (!(l > h)) {
                                R lap = kVisc * (h - l);
                                fv = v3<R>(((v2.x - v1.x) / (r1 * r2)) * lap * (R)mu,
                                           ((v2.y - v1.y) / (r1 * r2)) * lap * (R)mu,
                                           ((v2.z - v1.z) / (r1 * r2)) * lap * (R)mu);
                            }
                        }
                        part += fv;
                        if (mag_out) mag += len(fp) + len(fv);
                    }
                }
                internal += part;
            }
            total = internal * (R)m + g;
            V3<R> w = wall_terms<R>(pos, vel, total, cp);
            total += w;
            R speed = len(total);
            if (speed > vlim) {
                R t = vlim / speed;
                total = total * t;
            }
        }
        const uint32_t o = id[i];
        V3<R> dv = total * dt;
        v_out[4 * o + 0] = (R)vel.x + dv.x;
        v_out[4 * o + 1] = (R)vel.y + dv.y;
        v_out[4 * o + 2] = (R)vel.z + dv.z;
        v_out[4 * o + 3] = (R)vel.w + (R)0;
        if (acc_out) { acc_out[3 * i] = total.x; acc_out[3 * i + 1] = total.y; acc_out[3 * i + 2] = total.z; }
        if (mag_out) mag_out[i] = mag * (R)m;
    }
}

extern "C" {

void orc_forces_f32(const fg_float4* sp, const fg_float4* sv, const float* rho_p, const uint32_t* start,
                    const uint32_t* end, float* v_out, const uint32_t* id, uint32_t n, float m, float mu,
                    const fg_common_fluid_params* cp, const fg_kernels_cache* kc, int mode,
                    float* acc_out, float* mag_out) {
    forces_t<float, float>(sp, sv, rho_p, start, end, v_out, id, n, m, mu, *cp, *kc, mode, acc_out, mag_out);
}

void orc_forces_f64(const fg_float4* sp, const fg_float4* sv, const double* rho_p, const uint32_t* start,
                    const uint32_t* end, double* v_out, const uint32_t* id, uint32_t n, float m, float mu,
                    const fg_common_fluid_params* cp, const fg_kernels_cache* kc, int mode,
                    double* acc_out, double* mag_out) {
    forces_t<double, double>(sp, sv, rho_p, start, end, v_out, id, n, m, mu, *cp, *kc, mode, acc_out, mag_out);
}

/* ---- integrate (K_FluidSim.cu:418-438) --------------------------------- */
void orc_integrate(fg_float4* pos, const fg_float4* vel, uint32_t n, const fg_common_fluid_params* cp) {
    const float q = cp->deltaTimeIntegration / cp->simulation_scale;
    for (uint32_t i = 0; i < n; i++) {
        if (pos[i].w > 0.0f) {
            float ax = vel[i].x * q, ay = vel[i].y * q, az = vel[i].z * q, aw = vel[i].w * q;
            pos[i].x = pos[i].x + ax;
            pos[i].y = pos[i].y + ay;
            pos[i].z = pos[i].z + az;
            pos[i].w = pos[i].w + aw;
        }
    }
}

/* ---- one full step (HeterogeneousSimulator.cpp:54-127, 131-212) ---------
 * pos/vel are the unsorted state and are updated in place.  The scratch
 * arrays receive the intermediate results of the step (sizes: key,id n;
 * start,end numBuckets; sortedPos,sortedVel 4n floats; rho_p 2n floats).
 * rho_p must hold the previous step's values on entry (inactive particles
 * keep them). */
void orc_step(fg_float4* pos, fg_float4* vel, uint32_t n, const fg_fluid_params* fp,
              const fg_common_fluid_params* cp, const fg_kernels_cache* kc, int mode, uint32_t* key,
              uint32_t* id, uint32_t* start, uint32_t* end, fg_float4* sortedPos, fg_float4* sortedVel,
              float* rho_p) {
    const uint32_t nb = num_buckets(*cp);
    orc_hash(pos, n, nb, cp->cell_size, cp, mode, key, id);
    orc_sort_pairs(key, id, n);
    orc_cell_offsets(start, end, key, id, pos, sortedPos, vel, sortedVel, n, nb);
    orc_density_f32(sortedPos, rho_p, start, end, n, fp->gasConstant, fp->restDensity, fp->mass, cp, kc, mode);
    orc_forces_f32(sortedPos, sortedVel, rho_p, start, end, (float*)vel, id, n, fp->mass, fp->viscosity, cp,
                   kc, mode, nullptr, nullptr);
    orc_integrate(pos, vel, n, cp);
}

void orc_set_reverse_order(int r) { g_reverse = r; }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

} // extern "C"
