// fg_sph_mixed.hpp -- the mixed-precision density pass (LINEAR keys, fast
// math): candidates are culled two at a time in packed FP16 from 8-byte
// cell-relative offsets, and the survivors are evaluated exactly as the FP32
// pass evaluates every candidate, with a compensated (Neumaier) sum.
//
// Offsets.  For a particle in LINEAR cell (cx, cy, cz), with lower corner lo
// and u_a = clamp((x_a - lo_a) / cs_a, 0, 1) its position inside the cell
// (clamped: a particle outside the grid is projected onto its clamped cell),
// in units of the largest cell edge L (w_a = cs_a / L):
//     X = w_x (u_x + (cx mod 3) - 1.5),  Y = w_y (u_y - 0.5),  Z = w_z (u_z - 0.5).
// A stencil row holds the cells cx-1..cx+1 at (cy+dy, cz+dz).  For a
// candidate j of that row, Y_j - (Y_i - dy w_y) is the y distance of the two
// projections, likewise z.  X_j - X_i is their x distance up to a multiple
// of the period P = 3 w_x (the cells' indices are only known mod 3); the
// distance is min(|X_j - X_i|, P - |X_j - X_i|), exact whenever the true x
// distance is at most P/2 = 1.5 w_x, which every pair within h satisfies
// (h <= cs_x, the stencil's own requirement).
//
// Conservative cull.  Projection onto a box is 1-Lipschitz, so the projected
// distance never exceeds the true one.  The FP16 evaluation adds at most, in
// L^2 units, 8.4e-3 to r^2 for pairs with r <= h:
//   stored X, Y, Z: |.| <= 1.5, rounding <= 2^-11; the row shifts
//   (|.| <= 1.5): <= 2^-11; the differences (|.| < 3): <= 2^-10; P: <= 2^-10
//   (P - a is exact where it is taken, a in [P/2, P]); so each component is
//   off by at most e = 2^-9;
//   r^2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, -cut))): with |d| <= 1 the
//   components add at most 2 e (|dx| + |dy| + |dz|) <= 2 sqrt(3) e = 6.8e-3,
//   and the three roundings (results below 2) 3 * 2^-11 = 1.5e-3;
// plus the FP32 cell and offset arithmetic (below 1e-4).  The cut is
// (h/L)^2 + 2^-6 (1.6e-2), rounded up, so no pair that the FP32 test accepts
// is culled: the survivors' exact test then sets exactly the row-mask bits
// of the FP32 pass, and the force pass reads the same masks.
#pragma once
#include "fg_sph_tile.hpp"

namespace fg {

typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef short short2v __attribute__((ext_vector_type(2)));

struct HalfGeom {
    float ix, iy, iz;  // 1 / cell size
    float wx, wy, wz;  // cell size / L
    float per;         // 3 w_x
    float cut;         // (h/L)^2 + 2^-6, rounded up past FP16 rounding
};

__device__ __forceinline__ HalfGeom half_geom(const fg_common_fluid_params& cp) {
    HalfGeom g;
    const float L = fmaxf(cp.cell_size.x, fmaxf(cp.cell_size.y, cp.cell_size.z));
    g.ix = 1.0f / cp.cell_size.x;
    g.iy = 1.0f / cp.cell_size.y;
    g.iz = 1.0f / cp.cell_size.z;
    g.wx = cp.cell_size.x / L;
    g.wy = cp.cell_size.y / L;
    g.wz = cp.cell_size.z / L;
    g.per = 3.0f * g.wx;
    const float hc = (cp.kernel_radius / cp.simulation_scale) / L;
    g.cut = (hc * hc + 0.015625f) * 1.001953125f;  // x (1 + 2^-9) covers the FP16 rounding of the cut
    return g;
}

// The offsets (X, Y, Z) of a particle at p, in FP32.
__device__ __forceinline__ float3 half_offsets(const float4& p, const fg_common_fluid_params& cp,
                                               const HalfGeom& g) {
    const CellXYZ c = linear_cell(p.x, p.y, p.z, cp);
    const float lx = cp.volume_min.x + (float)c.x * cp.cell_size.x;
    const float ly = cp.volume_min.y + (float)c.y * cp.cell_size.y;
    const float lz = cp.volume_min.z + (float)c.z * cp.cell_size.z;
    // fmaxf first: a NaN coordinate becomes 0 (its exact tests fail anyway)
    const float ux = fminf(fmaxf((p.x - lx) * g.ix, 0.0f), 1.0f);
    const float uy = fminf(fmaxf((p.y - ly) * g.iy, 0.0f), 1.0f);
    const float uz = fminf(fmaxf((p.z - lz) * g.iz, 0.0f), 1.0f);
    return make_float3(g.wx * (ux + (float)(c.x % 3) - 1.5f), g.wy * (uy - 0.5f), g.wz * (uz - 0.5f));
}

__device__ __forceinline__ uint32_t pack_half2(float lo, float hi) {
    const half2v v = {(_Float16)lo, (_Float16)hi};
    return __builtin_bit_cast(uint32_t, v);
}

// Pair record r of the sorted particles: (X_r, X_r+1), (Y_r, Y_r+1),
// (Z_r, Z_r+1), 0 as packed FP16; records n and n+1 are zero (reads past a
// row's end are masked off).
__device__ __forceinline__ uint4* nbr_half(unsigned long long* masks, uint32_t n) {
    return reinterpret_cast<uint4*>(reinterpret_cast<char*>(masks) + nbr_half_offset(n));
}
__device__ __forceinline__ const uint4* nbr_half(const unsigned long long* masks, uint32_t n) {
    return reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(masks) + nbr_half_offset(n));
}

#pragma clang fp contract(off)
// r^2 - cut of the two candidates of a pair record, packed; a candidate
// survives when its half is negative.
__device__ __forceinline__ uint32_t cull_pair(const uint4& r, half2v tx, half2v ty, half2v tz, half2v per,
                                              half2v ncut) {
    const half2v X = __builtin_bit_cast(half2v, r.x);
    const half2v Y = __builtin_bit_cast(half2v, r.y);
    const half2v Z = __builtin_bit_cast(half2v, r.z);
    const half2v a = __builtin_elementwise_abs(X - tx);
    const half2v dx = __builtin_elementwise_min(a, per - a);
    const half2v dy = Y - ty, dz = Z - tz;
    const half2v d = __builtin_elementwise_fma(dz, dz, __builtin_elementwise_fma(dy, dy,
                                               __builtin_elementwise_fma(dx, dx, ncut)));
    return __builtin_bit_cast(uint32_t, d);
}

// Bits 0..3: candidates j..j+3 survive (records j and j+2).
__device__ __forceinline__ uint32_t cull_quad(const uint4& r0, const uint4& r1, half2v tx, half2v ty, half2v tz,
                                              half2v per, half2v ncut) {
    const short2v m0 = __builtin_bit_cast(short2v, cull_pair(r0, tx, ty, tz, per, ncut)) >> (short)15;
    const short2v m1 = __builtin_bit_cast(short2v, cull_pair(r1, tx, ty, tz, per, ncut)) >> (short)15;
    const uint32_t x = (__builtin_bit_cast(uint32_t, m0) & 0x00020001u) |
                       (__builtin_bit_cast(uint32_t, m1) & 0x00080004u);
    return (x & 0xFu) | (x >> 16);
}

// |p - q|^2 with the operations of density4 (the FP32 pass's test).
__device__ __forceinline__ float exact_r2(const float4& p, const float3& q) {
    const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}
__device__ __forceinline__ float poly6_term(float h2u, float r2) {
    const float t = h2u - r2;
    return (t * t) * t;
}

// sum + comp += x (Neumaier's compensated summation).
__device__ __forceinline__ void neumaier_add(float& sum, float& comp, float x) {
    const float t = sum + x;
    comp += (fabsf(sum) >= fabsf(x)) ? ((sum - t) + x) : ((x - t) + sum);
    sum = t;
}
#pragma clang fp contract(fast)

// One stencil row of the mixed density pass for one particle: the cull of
// the row's candidates from their FP16 pair records, then the exact FP32
// test and compensated sum of the survivors.  H and P are the pair records
// and sorted positions from the row's first candidate, in memory or (LDS)
// staged.
struct MixedRow {
    float4 p;
    half2v tx, per, ncut;
    float hhu, h2u;
    float sum, comp;

    // The survivors among candidates P[bits of word], four per iteration
    // with their loads issued together; returns the bits of those within h
    // (the FP32 test of density4).
    template <bool LDS>
    __device__ __forceinline__ uint32_t exact(uint32_t word, const float4* __restrict__ P) {
        auto ld = [](const float4* q) { return LDS ? ld_xyz_lds(q) : ld_xyz(q); };
        uint32_t hits = 0u;
        while (word) {
            const uint32_t k0 = (uint32_t)__builtin_ctz(word);
            word &= word - 1u;
            const bool v1 = word != 0u;
            const uint32_t k1 = v1 ? (uint32_t)__builtin_ctz(word) : k0;
            word &= word - 1u;
            const bool v2 = word != 0u;
            const uint32_t k2 = v2 ? (uint32_t)__builtin_ctz(word) : k0;
            word &= word - 1u;
            const bool v3 = word != 0u;
            const uint32_t k3 = v3 ? (uint32_t)__builtin_ctz(word) : k0;
            word &= word - 1u;
            const float3 q0 = ld(P + k0), q1 = ld(P + k1);
            const float3 q2 = ld(P + k2), q3 = ld(P + k3);
            const float r0 = exact_r2(p, q0), r1 = exact_r2(p, q1);
            const float r2 = exact_r2(p, q2), r3 = exact_r2(p, q3);
            if (r0 < hhu) {
                neumaier_add(sum, comp, poly6_term(h2u, r0));
                hits |= 1u << k0;
            }
            if (v1 && r1 < hhu) {
                neumaier_add(sum, comp, poly6_term(h2u, r1));
                hits |= 1u << k1;
            }
            if (v2 && r2 < hhu) {
                neumaier_add(sum, comp, poly6_term(h2u, r2));
                hits |= 1u << k2;
            }
            if (v3 && r3 < hhu) {
                neumaier_add(sum, comp, poly6_term(h2u, r3));
                hits |= 1u << k3;
            }
        }
        return hits;
    }

    // Survivor bits of the cnt (<= 32) candidates from H[0].  (Keeping the
    // next group's records in flight, eight candidates per iteration, needs
    // 76 VGPRs and was no faster: the loop is bound by issue, not latency.)
    __device__ __forceinline__ uint32_t cull_word(const uint4* __restrict__ H, uint32_t cnt, half2v ty,
                                                  half2v tz) const {
        uint32_t word = 0u;
        for (uint32_t t = 0; t < cnt; t += 4) {
            const uint4 r0 = H[t], r1 = H[t + 2];
            word |= cull_quad(r0, r1, tx, ty, tz, per, ncut) << t;
        }
        return cnt < 32u ? word & ((1u << cnt) - 1u) : word;
    }

    // The row's mask: 64 candidates at a time; a row longer than 64 is never
    // read through its mask (the force pass re-tests it), which is 0.
    template <bool LDS>
    __device__ __forceinline__ unsigned long long row(const uint4* __restrict__ H, const float4* __restrict__ P,
                                                      uint32_t len, half2v ty, half2v tz) {
        unsigned long long msk = 0ull;
        for (uint32_t c0 = 0; c0 < len; c0 += 64u) {
            const uint32_t l = min(len - c0, 64u);
            const uint32_t w0 = cull_word(H + c0, min(l, 32u), ty, tz);
            const uint32_t w1 = l > 32u ? cull_word(H + c0 + 32u, l - 32u, ty, tz) : 0u;
            const uint32_t e0 = exact<LDS>(w0, P + c0);
            const uint32_t e1 = exact<LDS>(w1, P + c0 + 32u);
            msk = (unsigned long long)e0 | ((unsigned long long)e1 << 32);
        }
        return len > kRowMaskBits ? 0ull : msk;
    }
};

// Density of sorted particle i (i < n) in the mixed form: the row masks and
// the neighbour record of the FP32 pass (density_particle<true>), and
// rho = m kPoly6 s^6 sum (h^2 - r^2)^3 over the same pairs, summed in
// ascending sorted order within rows in order, compensated.
//
// Given an LDS slice (`stage`, kMixedStage pair records then kMixedStage
// positions) and a full wavefront, each row's run of pair records and
// positions is staged there first, as in density_rows_staged; every lane
// then takes part (an inactive particle's lane with no rows).
constexpr uint32_t kMixedStage = 128;
__device__ __forceinline__ void density_particle_mixed(uint32_t i, const float4* __restrict__ sp,
                                                       const float4* __restrict__ sv, float2* __restrict__ rho_p,
                                                       const uint32_t* __restrict__ start,
                                                       const uint32_t* __restrict__ end, uint32_t n, float k,
                                                       float rho0, float m, const fg_common_fluid_params& cp,
                                                       const fg_kernels_cache& kc,
                                                       unsigned long long* __restrict__ masks,
                                                       float4* __restrict__ stage = nullptr) {
    const float4 p = sp[i];
    float4* rec = nbr_records(masks, n);
    const float4 v = sv[i];
    const bool active = p.w > 0.0f;
    const bool staged_rows = stage != nullptr && __ballot(1) == ~0ull;
    if (!active && !staged_rows) {
        store_record(i, rho_p[i], p, v, rec);
        return;
    }
    const float s = cp.simulation_scale, s2 = s * s;
    const float h = cp.kernel_radius, hh = h * h;
    const float hu = h / s, hhu = hu * hu, h2u = cp.kernel_radius_squared / s2;
    const float cut = hh * 1.0001f;
    const RowWalker w = make_rows(p, cp);
    const HalfGeom g = half_geom(cp);
    const float3 o = half_offsets(p, cp, g);
    const uint4* __restrict__ hr = nbr_half(masks, n);
    MixedRow M{p,
               half2v{(_Float16)o.x, (_Float16)o.x},
               half2v{(_Float16)g.per, (_Float16)g.per},
               half2v{(_Float16)(-g.cut), (_Float16)(-g.cut)},
               hhu,
               h2u,
               0.0f,
               0.0f};
    uint4* hst = reinterpret_cast<uint4*>(stage);
    const float4* pst = stage + (kMixedStage + 4u);

    for (int row = 0; row < 9; row++) {
        const int dz = row / 3 - 1, dy = row % 3 - 1;
        uint32_t rb = 0, re = 0;
        const bool has = active && row_range(w, dy, dz, cut, cp, 0u, start, end, n, rb, re);
        uint32_t lo = 0, hi = 0;
        const bool staged = staged_rows && wave_run(has, rb, re, kMixedStage, lo, hi);  // the same on every lane
        if (staged) {
            // the pair records of the run and the two after it (a quad of
            // candidates reads the records t and t + 2), and the positions
            for (uint32_t q = __lane_id(); q < hi - lo + 2u; q += 64u) hst[q] = hr[lo + q];
            for (uint32_t q = __lane_id(); q < hi - lo; q += 64u) stage[kMixedStage + 4u + q] = sp[lo + q];
            wave_lds_sync();
        }
        unsigned long long msk = 0ull;
        if (has) {
            const half2v ty = {(_Float16)(o.y - (float)dy * g.wy), (_Float16)(o.y - (float)dy * g.wy)};
            const half2v tz = {(_Float16)(o.z - (float)dz * g.wz), (_Float16)(o.z - (float)dz * g.wz)};
            msk = staged ? M.row<true>(hst + (rb - lo), pst + (rb - lo), re - rb, ty, tz)
                         : M.row<false>(hr + rb, sp + rb, re - rb, ty, tz);
        }
        if (active) masks[(size_t)row * n + i] = msk;
        if (staged) wave_lds_sync();  // every lane's reads of this row before the next row's stores
    }
    if (active)
        store_density(i, m * kc.kPoly6 * (s2 * s2 * s2) * (M.sum + M.comp), k, rho0, rho_p, rec, p, v);
    else
        store_record(i, rho_p[i], p, v, rec);
}

}  // namespace fg
