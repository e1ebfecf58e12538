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

/* ======================================================================
 * Rigid-body coupling
 * ====================================================================== */

/* A bucket grid walked by the penalty-contact stencil: the fluid grid (its
 * key mode), or a body-sphere grid (REF_XOR). */
struct OGrid {
    int mode;
    fg_float3 cs;
    uint32_t nb;
    fg_float3 vmin;
    fg_int3 dims;
};

} // extern "C"

namespace {

inline uint32_t ogrid_key_of_cell(const OGrid& g, int cx, int cy, int cz) {
    return (uint32_t)cx + (uint32_t)g.dims.x * ((uint32_t)cy + (uint32_t)g.dims.y * (uint32_t)cz);
}

/* K_DeviceFunctions.cu:32-106 (ProcessCellForcesD) */
template <class R>
V3<R> process_cell(const fg_float4& sp, const fg_float4& sv, R sr, R cr, const fg_float4* cpos, const fg_float4* cvel,
                   uint32_t key, const uint32_t* start, const uint32_t* end, R ks, R kd, bool w_index, bool group,
                   int w_group, R eps) {
    V3<R> F = v3<R>(0, 0, 0);
    const R mind = sr + cr + eps;
    if (start[key] != kEmpty) {
        for (uint32_t i = start[key]; i < end[key]; i++) {
            const fg_float4 pos = cpos[i];
            fg_float4 vel{0.0f, 0.0f, 0.0f, 0.0f};
            if (group && f2i_sat(pos.w) == w_group) continue;
            if (cvel) vel = w_index ? cvel[f2i_sat(pos.w)] : cvel[i];
            const V3<R> rel = xyz<R>(pos) - xyz<R>(sp);
            const R lg = len(rel);
            if (lg < mind) {
                const V3<R> n = rel / lg;
                const V3<R> rv = xyz<R>(vel) - xyz<R>(sv);
                F += repulsion(rv, n, lg, kd, ks);
            }
        }
    }
    return F;
}

/* K_DeviceFunctions.cu:113-159 (ProcessCollisionForcesD); LINEAR grids use
 * the cell-offset stencil with out-of-range cells skipped. */
template <class R>
V3<R> process_collision(const fg_float4& sp, const fg_float4& sv, R sr, const fg_float4* cpos, const fg_float4* cvel,
                        R cr, R ks, R kd, fg_int3 rel, const uint32_t* start, const uint32_t* end, const OGrid& g,
                        bool w_index, bool group, int w_group, R eps) {
    V3<R> F = v3<R>(0, 0, 0);
    int cx = 0, cy = 0, cz = 0;
    if (g.mode == FG_KEY_LINEAR) {
        cx = cell_coord(sp.x, g.vmin.x, g.cs.x, g.dims.x);
        cy = cell_coord(sp.y, g.vmin.y, g.cs.y, g.dims.y);
        cz = cell_coord(sp.z, g.vmin.z, g.cs.z, g.dims.z);
    }
    for (int z = -rel.z; z <= rel.z; z++)
        for (int y = -rel.y; y <= rel.y; y++)
            for (int x = -rel.x; x <= rel.x; x++) {
                uint32_t key;
                if (g.mode == FG_KEY_REF_XOR) {
                    fg_float3 off = g.cs;
                    off.x *= (float)x; off.y *= (float)y; off.z *= (float)z;
                    key = hash_ref_xor(sp.x + off.x, sp.y + off.y, sp.z + off.z, g.cs, g.nb);
                } else {
                    const int a = cx + x, b = cy + y, c = cz + z;
                    if (a < 0 || b < 0 || c < 0 || a >= g.dims.x || b >= g.dims.y || c >= g.dims.z) continue;
                    key = ogrid_key_of_cell(g, a, b, c);
                }
                F += process_cell<R>(sp, sv, sr, cr, cpos, cvel, key, start, end, ks, kd, w_index, group, w_group,
                                     eps);
            }
    return F;
}

/* Static-body grid of the coupled force pass: sorted sphere centres, cell
 * ranges, grid and sphere radius (no velocities). */
struct OStatic {
    const fg_float4* pos;
    const uint32_t* start;
    const uint32_t* end;
    OGrid grid;
    float radius;
};

/* Coupled fluid forces, K_FluidSim.cu:315-413: the static grid (st, may be
 * null; contact against spheres with no velocity) is added to the external
 * force before the dynamic grid (rb_pos may be null). */
template <class R, class In>
void forces_coupled_t(const fg_float4* sp, const fg_float4* sv, const In* rho_p, const uint32_t* start,
                      const uint32_t* end, R* v_out, const uint32_t* id, uint32_t n, float m, float mu,
                      const fg_common_fluid_params& cp, const fg_kernels_cache& kc, int mode, float fluid_radius,
                      fg_int3 relation, float body_radius, float ks, float kd, const fg_float4* rb_pos,
                      const fg_float4* rb_vel, const uint32_t* rb_start, const uint32_t* rb_end, const OGrid& rbg,
                      const OStatic* st, R* acc_out) {
    const R scale = (R)cp.simulation_scale, h = (R)cp.kernel_radius;
    const R kSpiky = (R)kc.kSpikyGradient, kVisc = (R)kc.kViscosityLaplacian;
    const R vlim = (R)cp.velocity_limit, dt = (R)cp.deltaTimeIntegration;
    const V3<R> g = v3<R>(cp.global_external_force.x, cp.global_external_force.y, cp.global_external_force.z);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const uint32_t i = (uint32_t)ii;
        const fg_float4 pos = sp[i], vel = sv[i];
        V3<R> total = v3<R>(0, 0, 0);
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
                    for (uint32_t j = s; j < e; j++) {
                        if (j == i) continue;
                        V3<R> p2 = xyz<R>(sp[j]) * scale;
                        const R r2 = (R)rho_p[2 * j], pr2 = (R)rho_p[2 * j + 1];
                        const V3<R> v2 = xyz<R>(sv[j]);
                        V3<R> diff = p1 - p2;
                        R l = len(diff);
                        if (!(l > h)) {
                            R pp = (pr1 + pr2) / (r1 * r2);
                            part += (diff / l) * (kSpiky * ((h - l) * (h - l))) * pp * (R)-1.0f;
                            R lap = kVisc * (h - l);
                            part += v3<R>(((v2.x - v1.x) / (r1 * r2)) * lap * (R)mu,
                                          ((v2.y - v1.y) / (r1 * r2)) * lap * (R)mu,
                                          ((v2.z - v1.z) / (r1 * r2)) * lap * (R)mu);
                        }
                    }
                }
                internal += part;
            }
            V3<R> external = v3<R>(0, 0, 0);
            if (st)
                external += process_collision<R>(pos, vel, (R)fluid_radius, st->pos, nullptr, (R)st->radius, (R)ks,
                                                 (R)kd, relation, st->start, st->end, st->grid, false, false, 0,
                                                 (R)0.003f);
            if (rb_pos)
                external += process_collision<R>(pos, vel, (R)fluid_radius, rb_pos, rb_vel, (R)body_radius, (R)ks,
                                                 (R)kd, relation, rb_start, rb_end, rbg, true, false, 0,
                                                 (R)0.003f);
            total = internal * (R)m + g + external;
            V3<R> w = wall_terms<R>(pos, vel, total, cp);
            total += w;
            R speed = len(total);
            if (speed > vlim) total = total * (vlim / speed);
        }
        const uint32_t o = id[i];
        V3<R> dv = total * dt;
        v_out[4 * o + 0] = (R)vel.x + dv.x;
        v_out[4 * o + 1] = (R)vel.y + dv.y;
        v_out[4 * o + 2] = (R)vel.z + dv.z;
        v_out[4 * o + 3] = (R)vel.w + (R)0;
        if (acc_out) { acc_out[3 * i] = total.x; acc_out[3 * i + 1] = total.y; acc_out[3 * i + 2] = total.z; }
    }
}

} // namespace

extern "C" {

/* K_HashGrid.cu:43-63: world = conj(q) (p,0) q + cm[w]; w kept; key; id. */
void orc_body_transform(const fg_float4* rel, const fg_float4* cm, const fg_float4* q, uint32_t m, const OGrid* g,
                        fg_float4* world, uint32_t* key, uint32_t* id) {
    for (uint32_t i = 0; i < m; i++) {
        const fg_float4 p = rel[i];
        const int b = f2i_sat(p.w);
        const fg_float4 o = q[b];
        /* Quaternion_ProductD (K_MathUtils.cu:160-169) twice, conj(q) first. */
        auto prod = [](fg_float4 a, fg_float4 c) {
            float w = a.w * c.w - (a.x * c.x + a.y * c.y + a.z * c.z);
            float cx = (a.y * c.z) - (a.z * c.y), cy = (a.z * c.x) - (a.x * c.z), cz = (a.x * c.y) - (a.y * c.x);
            return fg_float4{(a.w * c.x) + (c.w * a.x) + cx, (a.w * c.y) + (c.w * a.y) + cy,
                             (a.w * c.z) + (c.w * a.z) + cz, w};
        };
        fg_float4 pp{p.x, p.y, p.z, 0.0f};
        fg_float4 t = prod(prod(fg_float4{-o.x, -o.y, -o.z, o.w}, pp), o);
        fg_float4 a{t.x + cm[b].x, t.y + cm[b].y, t.z + cm[b].z, t.w + cm[b].w};
        a.w = p.w;
        world[i] = a;
        id[i] = i;
        key[i] = (g->mode == FG_KEY_REF_XOR)
                     ? hash_ref_xor(a.x, a.y, a.z, g->cs, g->nb)
                     : ogrid_key_of_cell(*g, cell_coord(a.x, g->vmin.x, g->cs.x, g->dims.x),
                                         cell_coord(a.y, g->vmin.y, g->cs.y, g->dims.y),
                                         cell_coord(a.z, g->vmin.z, g->cs.z, g->dims.z));
    }
}

void orc_forces_coupled_f32(const fg_float4* sp, const fg_float4* sv, const float* rho_p, const uint32_t* start,
                            const uint32_t* end, float* v_out, const uint32_t* id, uint32_t n, float m, float mu,
                            const fg_common_fluid_params* cp, const fg_kernels_cache* kc, int mode,
                            float fluid_radius, fg_int3 relation, float body_radius, float ks, float kd,
                            const fg_float4* rb_pos, const fg_float4* rb_vel, const uint32_t* rb_start,
                            const uint32_t* rb_end, const OGrid* rbg, const OStatic* st, float* acc_out) {
    forces_coupled_t<float, float>(sp, sv, rho_p, start, end, v_out, id, n, m, mu, *cp, *kc, mode, fluid_radius,
                                   relation, body_radius, ks, kd, rb_pos, rb_vel, rb_start, rb_end, *rbg, st,
                                   acc_out);
}

void orc_forces_coupled_f64(const fg_float4* sp, const fg_float4* sv, const double* rho_p, const uint32_t* start,
                            const uint32_t* end, double* v_out, const uint32_t* id, uint32_t n, float m, float mu,
                            const fg_common_fluid_params* cp, const fg_kernels_cache* kc, int mode,
                            float fluid_radius, fg_int3 relation, float body_radius, float ks, float kd,
                            const fg_float4* rb_pos, const fg_float4* rb_vel, const uint32_t* rb_start,
                            const uint32_t* rb_end, const OGrid* rbg, const OStatic* st, double* acc_out) {
    forces_coupled_t<double, double>(sp, sv, rho_p, start, end, v_out, id, n, m, mu, *cp, *kc, mode, fluid_radius,
                                     relation, body_radius, ks, kd, rb_pos, rb_vel, rb_start, rb_end, *rbg, st,
                                     acc_out);
}

/* K_RigidBodySim.cu:66-99: per body sphere, penalty force from the fluid and
 * torque (x - cm[w]) x F; F.w = torque.w = (float)(int)w. */
void orc_body_forces_from_fluid(const fg_float4* rb_pos, const fg_float4* rb_vel, const fg_float4* rb_cm,
                                fg_float4* out_f, fg_float4* out_t, uint32_t m, float radius,
                                const fg_float4* sph_pos, const fg_float4* sph_vel, const uint32_t* sph_start,
                                const uint32_t* sph_end, float sph_radius, const OGrid* fg, fg_int3 relation,
                                float ks, float kd) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)m; ii++) {
        const uint32_t i = (uint32_t)ii;
        const fg_float4 pos = rb_pos[i];
        const int b = f2i_sat(pos.w);
        const fg_float4 vel = rb_vel[b], cm = rb_cm[b];
        V3<float> F = process_collision<float>(pos, vel, radius, sph_pos, sph_vel, sph_radius, ks, kd, relation,
                                               sph_start, sph_end, *fg, false, false, 0, 0.0001f);
        out_f[i] = fg_float4{F.x, F.y, F.z, (float)b};
        const fg_float4 dir{pos.x - cm.x, pos.y - cm.y, pos.z - cm.z, pos.w - cm.w};
        out_t[i] = fg_float4{(dir.y * F.z) - (dir.z * F.y), (dir.z * F.x) - (dir.x * F.z),
                             (dir.x * F.y) - (dir.y * F.x), (float)b};
    }
}

/* K_RigidBodySim.cu:18-61: per body sphere, penalty force from the spheres of
 * the other bodies in the dynamic grid (same-body spheres skipped by w,
 * velocities by w) and from the static grid (no velocities), margin 0.001,
 * stencil (1,1,1); torque (x - cm[w]) x F.  dyn_pos or st may be null.  With
 * accumulate the force and torque are added to out_f/out_t (the fluid's
 * part), in float, else written; w = (float)(int)w. */
void orc_body_forces_from_bodies(const fg_float4* rb_pos, const fg_float4* rb_vel, const fg_float4* rb_cm,
                                 fg_float4* out_f, fg_float4* out_t, uint32_t m, float radius,
                                 const fg_float4* dyn_pos, const uint32_t* dyn_start, const uint32_t* dyn_end,
                                 const OGrid* dyn_grid, float dyn_radius, const OStatic* st, float ks, float kd,
                                 int accumulate) {
    const fg_int3 rel{1, 1, 1};
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)m; ii++) {
        const uint32_t i = (uint32_t)ii;
        const fg_float4 pos = rb_pos[i];
        const int b = f2i_sat(pos.w);
        const fg_float4 vel = rb_vel[b], cm = rb_cm[b];
        V3<float> F = v3<float>(0, 0, 0);
        if (dyn_pos)
            F += process_collision<float>(pos, vel, radius, dyn_pos, rb_vel, dyn_radius, ks, kd, rel, dyn_start,
                                          dyn_end, *dyn_grid, true, true, b, 0.001f);
        if (st)
            F += process_collision<float>(pos, vel, radius, st->pos, nullptr, st->radius, ks, kd, rel, st->start,
                                          st->end, st->grid, false, false, 0, 0.001f);
        const fg_float4 dir{pos.x - cm.x, pos.y - cm.y, pos.z - cm.z, pos.w - cm.w};
        const fg_float4 t{(dir.y * F.z) - (dir.z * F.y), (dir.z * F.x) - (dir.x * F.z), (dir.x * F.y) - (dir.y * F.x),
                          (float)b};
        if (accumulate) {
            out_f[i] = fg_float4{out_f[i].x + F.x, out_f[i].y + F.y, out_f[i].z + F.z, (float)b};
            out_t[i] = fg_float4{out_t[i].x + t.x, out_t[i].y + t.y, out_t[i].z + t.z, (float)b};
        } else {
            out_f[i] = fg_float4{F.x, F.y, F.z, (float)b};
            out_t[i] = t;
        }
    }
}

/* Per-body sums in FP64 of per-sphere forces and torques (the reduction of
 * K_RigidBodySim.cu:102-155 and RigidBodySimulator.cpp:522-566, whose CUDPP
 * scan sums in FP32); body index from F.w, sums in sphere order. */
void orc_body_reduce_f64(const fg_float4* f, const fg_float4* t, uint32_t m, uint32_t nbodies, double* out_f,
                         double* out_t) {
    std::fill(out_f, out_f + 3 * nbodies, 0.0);
    std::fill(out_t, out_t + 3 * nbodies, 0.0);
    for (uint32_t i = 0; i < m; i++) {
        const int b = f2i_sat(f[i].w);
        if (b < 0 || (uint32_t)b >= nbodies) continue;
        out_f[3 * b] += f[i].x; out_f[3 * b + 1] += f[i].y; out_f[3 * b + 2] += f[i].z;
        out_t[3 * b] += t[i].x; out_t[3 * b + 1] += t[i].y; out_t[3 * b + 2] += t[i].z;
    }
}

/* K_RigidBodySim.cu:160-262 (IntegrateRBK) with K_MathUtils.cu:72-236. */
void orc_integrate_rb(fg_float4* pos, const float* mass, fg_float4* vel, const fg_float4* forces, uint32_t nb,
                      float dt, fg_float4 ext, fg_float3 vmin, fg_float3 vmax, fg_float4* L, const fg_float4* torque,
                      const fg_float4* inv_inertia, fg_float4* orient, float vlimit) {
    for (uint32_t idx = 0; idx < nb; idx++) {
        fg_float4 F = forces[idx];
        const float spd = std::sqrt(F.x * F.x + F.y * F.y + F.z * F.z + F.w * F.w);
        if (spd > vlimit) {
            const float t = vlimit / spd;
            F = fg_float4{F.x * t, F.y * t, F.z * t, F.w * t};
        }
        fg_float4 p = pos[idx];
        const float mm = mass[idx];
        const fg_float4 acc{((F.x + ext.x) / mm) * dt, ((F.y + ext.y) / mm) * dt, ((F.z + ext.z) / mm) * dt,
                            ((F.w + ext.w) / mm) * dt};
        fg_float4 v{vel[idx].x + (acc.x * dt), vel[idx].y + (acc.y * dt), vel[idx].z + (acc.z * dt),
                    vel[idx].w + (acc.w * dt)};
        p.x += v.x * dt;
        p.y += v.y * dt;
        p.z += v.z * dt;
        fg_float4 l = L[idx];
        const fg_float4 tq = torque[idx];
        l = fg_float4{l.x + tq.x * dt, l.y + tq.y * dt, l.z + tq.z * dt, l.w + tq.w * dt};
        l = fg_float4{l.x * 0.9f, l.y * 0.9f, l.z * 0.9f, l.w * 0.9f};
        L[idx] = l;
        /* AngularVelocityD: R I^-1 R^T L */
        const fg_float4 q = orient[idx];
        float R[3][3];
        R[0][0] = 1 - (2 * q.y * q.y) - (2 * q.z * q.z);
        R[0][1] = (2 * q.x * q.y) + (2 * q.w * q.z);
        R[0][2] = (2 * q.x * q.z) - (2 * q.w * q.y);
        R[1][0] = (2 * q.x * q.y) - (2 * q.w * q.z);
        R[1][1] = 1 - (2 * q.x * q.x) - (2 * q.z * q.z);
        R[1][2] = (2 * q.y * q.z) + (2 * q.w * q.x);
        R[2][0] = (2 * q.x * q.z) + (2 * q.w * q.y);
        R[2][1] = (2 * q.y * q.z) - (2 * q.w * q.x);
        R[2][2] = 1 - (2 * q.x * q.x) - (2 * q.y * q.y);
        const fg_float4* Ii = inv_inertia + 3 * idx;
        float I[3][3] = {{Ii[0].x, Ii[0].y, Ii[0].z}, {Ii[1].x, Ii[1].y, Ii[1].z}, {Ii[2].x, Ii[2].y, Ii[2].z}};
        float T[3][3], G[3][3];
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) T[a][c] = R[a][0] * I[0][c] + R[a][1] * I[1][c] + R[a][2] * I[2][c];
        for (int a = 0; a < 3; a++)  /* T * R^T */
            for (int c = 0; c < 3; c++) G[a][c] = T[a][0] * R[c][0] + T[a][1] * R[c][1] + T[a][2] * R[c][2];
        fg_float4 w{G[0][0] * l.x + G[0][1] * l.y + G[0][2] * l.z, G[1][0] * l.x + G[1][1] * l.y + G[1][2] * l.z,
                    G[2][0] * l.x + G[2][1] * l.y + G[2][2] * l.z, 0.0f};
        auto prod = [](fg_float4 a, fg_float4 c) {
            float ww = a.w * c.w - (a.x * c.x + a.y * c.y + a.z * c.z);
            float cx = (a.y * c.z) - (a.z * c.y), cy = (a.z * c.x) - (a.x * c.z), cz = (a.x * c.y) - (a.y * c.x);
            return fg_float4{(a.w * c.x) + (c.w * a.x) + cx, (a.w * c.y) + (c.w * a.y) + cy,
                             (a.w * c.z) + (c.w * a.z) + cz, ww};
        };
        fg_float4 wq = prod(w, q);
        wq = fg_float4{wq.x * 0.5f, wq.y * 0.5f, wq.z * 0.5f, wq.w * 0.5f};
        const fg_float4 scaled{wq.x * dt * 0.1f, wq.y * dt * 0.1f, wq.z * dt * 0.1f, wq.w * dt * 0.1f};
        fg_float4 dq = prod(scaled, q);
        fg_float4 nq{q.x + dq.x, q.y + dq.y, q.z + dq.z, q.w + dq.w};
        const float norm = std::sqrt(nq.w * nq.w + nq.x * nq.x + nq.y * nq.y + nq.z * nq.z);
        nq = fg_float4{nq.x / norm, nq.y / norm, nq.z / norm, nq.w / norm};
        orient[idx] = nq;
        if (p.y < vmin.y) { p.y = vmin.y; v.y *= -0.7f; }
        if (p.x <= vmin.x) { p.x = vmin.x; v.x *= -0.7f; }
        else if (p.x >= vmax.x) { p.x = vmax.x; v.x *= -0.7f; }
        if (p.z <= vmin.z) { p.z = vmin.z; v.z *= -0.7f; }
        else if (p.z >= vmax.z) { p.z = vmax.z; v.z *= -0.7f; }
        pos[idx] = p;
        vel[idx] = v;
    }
}

/* Centre of mass and inertia tensor about it (FP64) of equal point masses
 * M/count at the sphere centres of one body. */
void orc_body_mass_props_f64(const fg_float4* world, uint32_t count, double mass, double* cm3, double* inertia9) {
    double c[3] = {0, 0, 0};
    for (uint32_t i = 0; i < count; i++) { c[0] += world[i].x; c[1] += world[i].y; c[2] += world[i].z; }
    for (int a = 0; a < 3; a++) c[a] /= (double)count;
    double I[9] = {0};
    const double mi = mass / (double)count;
    for (uint32_t i = 0; i < count; i++) {
        const double r[3] = {world[i].x - c[0], world[i].y - c[1], world[i].z - c[2]};
        const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) I[3 * a + b] += mi * ((a == b ? r2 : 0.0) - r[a] * r[b]);
    }
    for (int a = 0; a < 3; a++) cm3[a] = c[a];
    for (int a = 0; a < 9; a++) inertia9[a] = I[a];
}

void orc_set_reverse_order(int r) { g_reverse = r; }

/* Generalized winding numbers in FP64 (Van Oosterom & Strackee 1983 solid
 * angle per triangle, summed over triangles / 4 pi): the checker of
 * fg_mesh_winding_lattice.  The reference has no counterpart (its body files,
 * Simulators/RigidBody.cpp:55-108, come from meshes sampled elsewhere).
 * Triangles naming a vertex >= nverts add 0. */
void orc_winding(const float* pts, uint32_t npts, const float* verts, uint32_t nverts, const uint32_t* tris,
                 uint32_t ntris, double* out) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)npts; i++) {
        const double px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
        double acc = 0.0;
        for (uint32_t t = 0; t < ntris; t++) {
            const uint32_t i0 = tris[3 * t], i1 = tris[3 * t + 1], i2 = tris[3 * t + 2];
            if (i0 >= nverts || i1 >= nverts || i2 >= nverts) continue;
            const double ax = verts[3 * i0] - px, ay = verts[3 * i0 + 1] - py, az = verts[3 * i0 + 2] - pz;
            const double bx = verts[3 * i1] - px, by = verts[3 * i1 + 1] - py, bz = verts[3 * i1 + 2] - pz;
            const double cx = verts[3 * i2] - px, cy = verts[3 * i2 + 1] - py, cz = verts[3 * i2 + 2] - pz;
            const double la = std::sqrt(ax * ax + ay * ay + az * az);
            const double lb = std::sqrt(bx * bx + by * by + bz * bz);
            const double lc = std::sqrt(cx * cx + cy * cy + cz * cz);
            const double det = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
            const double den = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + (bx * cx + by * cy + bz * cz) * la +
                               (cx * ax + cy * ay + cz * az) * lb;
            acc += 2.0 * std::atan2(det, den);
        }
        out[i] = acc / (4.0 * 3.14159265358979323846);
    }
}

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
