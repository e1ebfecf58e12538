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
#include "fg_walls.hpp"
#include "fg_contact.hpp"

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

// Acceleration after gravity, walls and clamp (K_FluidSim.cu:287-303), from
// the pair sum `internal`.
__device__ __forceinline__ F3 finish_accel(F3 internal, float m, const float4& pos, const float4& vel,
                                           const fg_common_fluid_params& cp, const Contacts& bc) {
    const F3 g = f3(cp.global_external_force.x, cp.global_external_force.y, cp.global_external_force.z);
    F3 total = add(mul(internal, m), g);
    // K_FluidSim.cu:389-390: body contact forces are added after the mass
    // factor (the uncoupled kernel adds nothing here).
    if (bc.coupled) total = add(total, fluid_external(pos, vel, bc));
    F3 wall = walls_no_penetration(xyz(pos), xyz(vel), cp);
    wall = add(wall, walls_no_slip(xyz(pos), xyz(vel), cp));
    total = add(total, wall);
    const float speed = len(total);
    if (speed > cp.velocity_limit) {
        const float t = cp.velocity_limit / speed;
        total = mul(total, t);
    }
    return total;
}

// vel[id] = v + a*dt, and optionally pos[id] = p + vel[id]*(dt/s) where
// p.w > 0 (K_FluidSim.cu:306-307 then :433-438).
__device__ __forceinline__ void write_back(uint32_t o, uint32_t n, const float4& pos, const float4& vel, F3 a,
                                           const fg_common_fluid_params& cp, float4* __restrict__ vout,
                                           float4* __restrict__ pout) {
    if (o >= n) return;
    const F3 dv = mul(a, cp.deltaTimeIntegration);
    const float4 nv = make_float4(vel.x + dv.x, vel.y + dv.y, vel.z + dv.z, vel.w + 0.0f);
    vout[o] = nv;
    if (pout && pos.w > 0.0f) {
        const float q = cp.deltaTimeIntegration / cp.simulation_scale;
        pout[o] = make_float4(pos.x + nv.x * q, pos.y + nv.y * q, pos.z + nv.z * q, pos.w + nv.w * q);
    }
}

// K_FluidSim.cu:260-309 (+ fused integrate when pout != nullptr)
template <int MODE>
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
        a = finish_accel(internal, m, pos, vel, cp, bc);
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
// Fast LINEAR kernels
// ---------------------------------------------------------------------------

// Per-particle view of the 3x3 stencil rows.  A row is the x-run of cells
// cx-1..cx+1 at (cy+dy, cz+dz); its particles are sorted[row.begin, row.end).
struct RowWalker {
    float gx_lo, gx_hi;  // squared scaled distance to the x-neighbour cells
    float gy[3], gz[3];  // squared scaled distance to the y/z neighbour slabs
    int cx, cy, cz;
};

__device__ __forceinline__ float sq(float v) { return v * v; }

__device__ __forceinline__ RowWalker make_rows(const float4& p, const fg_common_fluid_params& cp) {
    RowWalker w;
    const CellXYZ c = linear_cell(p.x, p.y, p.z, cp);
    w.cx = c.x;
    w.cy = c.y;
    w.cz = c.z;
    const float s = cp.simulation_scale;
    // Lower bounds of the distance from p to the neighbouring cells along
    // each axis (zero when p lies outside its own, clamped, cell).
    const float lx = cp.volume_min.x + (float)c.x * cp.cell_size.x, hx = lx + cp.cell_size.x;
    const float ly = cp.volume_min.y + (float)c.y * cp.cell_size.y, hy = ly + cp.cell_size.y;
    const float lz = cp.volume_min.z + (float)c.z * cp.cell_size.z, hz = lz + cp.cell_size.z;
    w.gx_lo = sq(fmaxf(0.0f, p.x - lx) * s);
    w.gx_hi = sq(fmaxf(0.0f, hx - p.x) * s);
    w.gy[0] = sq(fmaxf(0.0f, p.y - ly) * s);
    w.gy[1] = 0.0f;
    w.gy[2] = sq(fmaxf(0.0f, hy - p.y) * s);
    w.gz[0] = sq(fmaxf(0.0f, p.z - lz) * s);
    w.gz[1] = 0.0f;
    w.gz[2] = sq(fmaxf(0.0f, hz - p.z) * s);
    return w;
}

// Sorted range of row (dy,dz) after culling; returns false when empty.
// `cut` is h^2 enlarged by a small margin so that rounding in the cell
// bounds can never drop a pair with r < h.
__device__ __forceinline__ bool row_range(const RowWalker& w, int dy, int dz, float cut,
                                          const fg_common_fluid_params& cp, const uint32_t* __restrict__ start,
                                          const uint32_t* __restrict__ end, uint32_t n, uint32_t& rb,
                                          uint32_t& re) {
    const int cy = w.cy + dy, cz = w.cz + dz;
    if (cy < 0 || cz < 0 || cy >= cp.bucket_size.y || cz >= cp.bucket_size.z) return false;
    const float gyz = w.gy[dy + 1] + w.gz[dz + 1];
    if (gyz >= cut) return false;
    int x0 = w.cx - 1, x1 = w.cx + 1;
    if (x0 < 0 || gyz + w.gx_lo >= cut) x0 = w.cx;
    if (x1 >= cp.bucket_size.x || gyz + w.gx_hi >= cut) x1 = w.cx;
    const uint32_t kb = linear_key(0, cy, cz, cp);
    uint32_t b = kEmpty, e = 0;
    for (int x = x0; x <= x1; x++) {
        const uint32_t s0 = start[kb + (uint32_t)x];
        if (s0 != kEmpty) {
            if (b == kEmpty) b = s0;
            e = end[kb + (uint32_t)x];
        }
    }
    if (b == kEmpty) return false;
    rb = b;
    re = min(e, n);
    return rb < re;
}

// Blocks are handed to the eight XCDs round-robin; remapping them so that each
// XCD walks one contiguous slice of the sorted particles keeps the neighbour
// rows of consecutive blocks in the same L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nblocks) {
    constexpr uint32_t kXcd = 8;
    const uint32_t per = nblocks / kXcd, rem = nblocks % kXcd;
    const uint32_t x = b % kXcd, idx = b / kXcd;
    return x * per + min(x, rem) + idx;
}

// Neighbour bitmasks: for each of the 9 stencil rows of a particle, bit k of
// masks[row * n + i] is set when sorted particle rowBegin + k lies within h.
// Rows longer than 64 particles are not encoded and are re-tested.
constexpr uint32_t kRowMaskBits = 64;

// Poly6 sum over four candidates; bit u of the returned nibble is set when
// candidate u lies within h.
__device__ __forceinline__ uint32_t density4(const float4& p, const float4& q0, const float4& q1, const float4& q2,
                                             const float4& q3, float h2u, float hhu, float& acc) {
    const float4 q[4] = {q0, q1, q2, q3};
    uint32_t nib = 0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const float dx = p.x - q[u].x, dy = p.y - q[u].y, dz = p.z - q[u].z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float t = h2u - r2;
        const bool hit = r2 < hhu;
        acc = hit ? acc + t * t * t : acc;
        nib |= hit ? (1u << u) : 0u;
    }
    return nib;
}

template <bool MASKS>
__global__ void __launch_bounds__(kSphThreads)
    k_density_fast(const float4* __restrict__ sp, float2* __restrict__ rho_p, const uint32_t* __restrict__ start,
                   const uint32_t* __restrict__ end, uint32_t n, float k, float rho0, float m,
                   fg_common_fluid_params cp, fg_kernels_cache kc, unsigned long long* __restrict__ masks) {
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = sp[i];
    if (!(p.w > 0.0f)) return;
    // Distances are compared unscaled: r_u = r / s, so r < h <=> r_u^2 < (h/s)^2
    // and h^2 - r^2 = s^2 (h^2/s^2 - r_u^2).
    const float s = cp.simulation_scale, s2 = s * s;
    const float h = cp.kernel_radius, hh = h * h;
    const float hu = h / s, hhu = hu * hu, h2u = cp.kernel_radius_squared / s2;
    const float cut = hh * 1.0001f;
    const RowWalker w = make_rows(p, cp);
    float acc = 0.0f;
#pragma unroll
    for (int row = 0; row < 9; row++) {
        const int dz = row / 3 - 1, dy = row % 3 - 1;
        unsigned long long msk = 0ull;
        uint32_t rb, re;
        if (row_range(w, dy, dz, cut, cp, start, end, n, rb, re)) {
            const uint32_t len = re - rb, n4 = len & ~3u;
            const float4* pp = sp + rb;
            uint32_t kk = 0;
            if (n4) {
                // Software-pipelined groups of four: the next group's loads are
                // in flight while the current group is evaluated.
                float4 a0 = pp[0], a1 = pp[1], a2 = pp[2], a3 = pp[3];
                for (; kk + 4 < n4; kk += 4) {
                    pp += 4;
                    const float4 b0 = pp[0], b1 = pp[1], b2 = pp[2], b3 = pp[3];
                    const uint32_t nib = density4(p, a0, a1, a2, a3, h2u, hhu, acc);
                    if (MASKS) msk |= (unsigned long long)nib << (kk & 63u);
                    a0 = b0; a1 = b1; a2 = b2; a3 = b3;
                }
                const uint32_t nib = density4(p, a0, a1, a2, a3, h2u, hhu, acc);
                if (MASKS) msk |= (unsigned long long)nib << (kk & 63u);
                kk += 4;
                pp += 4;
            }
            if (kk < len) {  // 1-3 remaining candidates; the others are masked out
                const uint32_t rest = len - kk;
                const float4 far = make_float4(3.0e38f, 3.0e38f, 3.0e38f, 0.0f);
                const float4 t0 = pp[0];
                const float4 t1 = rest > 1 ? pp[1] : far;
                const float4 t2 = rest > 2 ? pp[2] : far;
                const uint32_t nib = density4(p, t0, t1, t2, far, h2u, hhu, acc);
                if (MASKS) msk |= (unsigned long long)nib << (kk & 63u);
            }
        }
        // Rows longer than 64 particles are never read through their mask.
        if (MASKS) masks[(size_t)row * n + i] = msk;
    }
    const float rho = m * kc.kPoly6 * (s2 * s2 * s2) * acc;
    rho_p[i] = make_float2(fmaxf(1.0f, rho), k * (rho - rho0));
}

// Pressure and viscosity terms of one neighbour j (r < h checked again, so
// the same code serves rows without a mask).
struct PairAcc {
    float px, py, pz, vx, vy, vz;
};

__device__ __forceinline__ void pair_term(PairAcc& A, uint32_t i, uint32_t j, const float4& pos, const float4& vel,
                                          float pri, float4 q, float4 vj, float2 dpj, float h, float s, float hhu) {
    const float dx = pos.x - q.x, dy = pos.y - q.y, dz = pos.z - q.z;
    const float r2u = dx * dx + dy * dy + dz * dz;  // unscaled
    if (r2u < hhu && j != i) {
        const float iru = __builtin_amdgcn_rsqf(r2u);
        const float hr = h - r2u * iru * s;
        const float irj = __builtin_amdgcn_rcpf(dpj.x);
        const float cp_ = hr * hr * (pri + dpj.y) * irj * iru;
        A.px += dx * cp_;
        A.py += dy * cp_;
        A.pz += dz * cp_;
        const float cv = hr * irj;
        A.vx += (vj.x - vel.x) * cv;
        A.vy += (vj.y - vel.y) * cv;
        A.vz += (vj.z - vel.z) * cv;
    }
}

template <bool MASKS>
__global__ void __launch_bounds__(kSphThreads)
    k_forces_fast(const float4* __restrict__ sp, const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                  const uint32_t* __restrict__ start, const uint32_t* __restrict__ end, float4* __restrict__ vout,
                  float4* __restrict__ pout, const uint32_t* __restrict__ id, uint32_t n, float m, float mu,
                  fg_common_fluid_params cp, fg_kernels_cache kc, const unsigned long long* __restrict__ masks,
                  Contacts bc) {
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 pos = sp[i];
    const float4 vel = sv[i];
    F3 a = f3(0.0f, 0.0f, 0.0f);
    if (pos.w > 0.0f) {
        const float h = cp.kernel_radius, hh = h * h;
        const float s = cp.simulation_scale;
        const float hu = h / s, hhu = hu * hu;
        const float cut = hh * 1.0001f;
        const float2 dpi = rho_p[i];
        const RowWalker w = make_rows(pos, cp);
        PairAcc A{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        for (int row = 0; row < 9; row++) {
            const int dz = row / 3 - 1, dy = row % 3 - 1;
            uint32_t rb, re;
            if (!row_range(w, dy, dz, cut, cp, start, end, n, rb, re)) continue;
            if (MASKS && re - rb <= kRowMaskBits) {
                unsigned long long mm = masks[(size_t)row * n + i];
                while (mm) {
                    const uint32_t k1 = (uint32_t)__builtin_ctzll(mm);
                    mm &= mm - 1ull;
                    const bool two = mm != 0ull;
                    const uint32_t k2 = two ? (uint32_t)__builtin_ctzll(mm) : k1;
                    if (two) mm &= mm - 1ull;
                    const uint32_t j1 = rb + k1, j2 = rb + k2;
                    const float4 q1 = sp[j1], q2 = sp[j2];
                    const float4 v1 = sv[j1], v2 = sv[j2];
                    const float2 d1 = rho_p[j1], d2 = rho_p[j2];
                    pair_term(A, i, j1, pos, vel, dpi.y, q1, v1, d1, h, s, hhu);
                    if (two) pair_term(A, i, j2, pos, vel, dpi.y, q2, v2, d2, h, s, hhu);
                }
            } else {
                for (uint32_t j = rb; j < re; j++) pair_term(A, i, j, pos, vel, dpi.y, sp[j], sv[j], rho_p[j], h, s, hhu);
            }
        }
        // pressure: (s d_u / r) = d_u / r_u, so no scale factor remains
        const float iri = 1.0f / dpi.x;
        const float cP = -kc.kSpikyGradient * iri;
        const float cV = kc.kViscosityLaplacian * mu * iri;
        const F3 internal = f3(cP * A.px + cV * A.vx, cP * A.py + cV * A.vy, cP * A.pz + cV * A.vz);
        a = finish_accel(internal, m, pos, vel, cp, bc);
    }
    write_back(id[i], n, pos, vel, a, cp, vout, pout);
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
static bool have_params(const char* who) {
    if (!state().have_common || !state().have_kernels) {
        set_error(FG_ERR_NOT_INITIALIZED, "%s: call gInitializeCommonFluidParams and gInitializeKernelParams first",
                  who);
        return false;
    }
    return true;
}

static bool use_fast() { return state().key_mode == FG_KEY_LINEAR && state().math_mode == FG_MATH_FAST; }

int launch_density(float4* sp, float2* rho_p, uint32_t* start, uint32_t* end, uint32_t n, float k, float rho0,
                   float m, unsigned long long* masks = nullptr) {
    const State& st = state();
    const dim3 grid(grid_for(n, kSphThreads)), block(kSphThreads);
    if (use_fast() && masks)
        hipLaunchKernelGGL(k_density_fast<true>, grid, block, 0, st.stream, sp, rho_p, start, end, n, k, rho0, m,
                           st.common, st.kernels, masks);
    else if (use_fast())
        hipLaunchKernelGGL(k_density_fast<false>, grid, block, 0, st.stream, sp, rho_p, start, end, n, k, rho0, m,
                           st.common, st.kernels, (unsigned long long*)nullptr);
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
    Contacts bc{};
    if (bcp) bc = *bcp;
    const State& st = state();
    const dim3 grid(grid_for(n, kSphThreads)), block(kSphThreads);
    if (use_fast() && masks)
        hipLaunchKernelGGL(k_forces_fast<true>, grid, block, 0, st.stream, sp, sv, rho_p, start, end, vout, pout, id,
                           n, m, mu, st.common, st.kernels, masks, bc);
    else if (use_fast())
        hipLaunchKernelGGL(k_forces_fast<false>, grid, block, 0, st.stream, sp, sv, rho_p, start, end, vout, pout, id,
                           n, m, mu, st.common, st.kernels, (const unsigned long long*)nullptr, bc);
    else if (st.key_mode == FG_KEY_LINEAR)
        hipLaunchKernelGGL(k_forces_ref<FG_KEY_LINEAR>, grid, block, 0, st.stream, sp, sv, rho_p, start, end, vout,
                           pout, id, n, m, mu, st.common, st.kernels, bc);
    else
        hipLaunchKernelGGL(k_forces_ref<FG_KEY_REF_XOR>, grid, block, 0, st.stream, sp, sv, rho_p, start, end, vout,
                           pout, id, n, m, mu, st.common, st.kernels, bc);
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
    return (size_t)9 * sizeof(unsigned long long) * (size_t)numParticles;
}

int fg_sph_density_nbr(fg_float4* orderedPositions, fg_float2* mass_pressure, unsigned int* gridStart,
                       unsigned int* gridEnd, unsigned long long* neighborMasks, unsigned int numParticles,
                       float gasConstant, float restDensity, float mass) {
    if (numParticles == 0) return FG_OK;
    if (!have_params("fg_sph_density_nbr")) return FG_ERR_NOT_INITIALIZED;
    if (!orderedPositions || !mass_pressure || !gridStart || !gridEnd || !neighborMasks) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sph_density_nbr: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    return launch_density(FG_F4(orderedPositions), FG_F2(mass_pressure), gridStart, gridEnd, numParticles,
                          gasConstant, restDensity, mass, neighborMasks);
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
