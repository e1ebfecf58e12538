// fg_contact.hpp -- penalty contact of one sphere against the spheres of a
// bucket grid (GPU_Kernels/K_DeviceFunctions.cu:32-159), written with the
// reference's arithmetic and contraction off.
#pragma once
#include "fg_internal.hpp"

namespace fg {
#pragma clang fp contract(off)

// One sphere of ProcessCellForcesD (K_DeviceFunctions.cu:32-106): sphere i
// at pos; for a sphere within mind = r_i + r_j + eps, F += (ks*l - kd*dot(n,
// v_j - v_i)) * n with n = (x_j - x_i)/l, l = |x_j - x_i| (the spring term
// uses the distance).  Velocities: per body (indexed by w) when w_index,
// else per sphere; none when vel is null.  A sphere's velocity enters only
// a contact term, so it is read only for spheres in contact.
//
// cut2 = mind^2 (1 + 2^-20), rounded, exceeds mind^2 (1 + 2^-21): a sphere
// with |x_j - x_i|^2 > cut2 has a distance above mind by more than two
// units in the last place of mind, so its rounded square root is not below
// mind and the exact test would reject it too.  Most spheres of a walked
// bucket are rejected by that comparison of the squared distance, without
// the square root.
__device__ __forceinline__ F3 contact_term(F3 F, const float4& pos, uint32_t i, const float4& sp, const float4& sv,
                                           float mind, float cut2, const float4* __restrict__ cvel, float ks,
                                           float kd, bool w_index, bool group, int w_group, uint32_t vlimit) {
    if (group && f2i_sat(pos.w) == w_group) return F;
    const F3 rel = sub(xyz(pos), xyz(sp));
    const float r2 = dot(rel, rel);
    if (!(r2 <= cut2)) return F;  // also NaN, which the exact test rejects as well
    const float lg = rn_sqrt(r2);  // len(rel)
    if (lg < mind) {
        float4 vel = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (cvel) {
            const uint32_t vi = w_index ? (uint32_t)f2i_sat(pos.w) : i;
            if (vi < vlimit) vel = cvel[vi];
        }
        const F3 n = divs(rel, lg);
        const F3 rv = sub(xyz(vel), xyz(sv));
        F = add(F, mul(ks * lg - kd * dot(n, rv), n));
    }
    return F;
}

// ProcessCellForcesD over the sorted range [s0, e_raw) of one bucket
// (spheres in order; two positions are loaded at a time).
__device__ __forceinline__ F3 contact_range(const float4& sp, const float4& sv, float sr, float cr,
                                            const float4* __restrict__ cpos, const float4* __restrict__ cvel,
                                            uint32_t s0, uint32_t e_raw, float ks, float kd, bool w_index,
                                            bool group, int w_group, float eps, uint32_t limit, uint32_t vlimit) {
    F3 F = f3(0.0f, 0.0f, 0.0f);
    const float mind = sr + cr + eps;
    const float cut2 = rn_mul(rn_mul(mind, mind), 0x1.00001p+0f);
    if (s0 != kEmpty) {
        const uint32_t e = min(e_raw, limit);
        uint32_t i = s0;
        for (; i + 1 < e; i += 2) {
            const float4 p0 = cpos[i], p1 = cpos[i + 1];
            F = contact_term(F, p0, i, sp, sv, mind, cut2, cvel, ks, kd, w_index, group, w_group, vlimit);
            F = contact_term(F, p1, i + 1, sp, sv, mind, cut2, cvel, ks, kd, w_index, group, w_group, vlimit);
        }
        if (i < e)
            F = contact_term(F, cpos[i], i, sp, sv, mind, cut2, cvel, ks, kd, w_index, group, w_group, vlimit);
    }
    return F;
}

// Bucket of stencil offset (x,y,z) around sp, or kEmpty for a LINEAR cell
// outside the grid.
__device__ __forceinline__ uint32_t contact_key(const float4& sp, const CellXYZ& c, int x, int y, int z,
                                                const GridDesc& g) {
    if (g.mode == FG_KEY_REF_XOR) {
        const float ox = g.cs.x * (float)x, oy = g.cs.y * (float)y, oz = g.cs.z * (float)z;
        return g.key_base + hash_ref_xor(sp.x + ox, sp.y + oy, sp.z + oz, g.cs.x, g.cs.y, g.cs.z, g.nb);
    }
    const int a = c.x + x, b = c.y + y, d = c.z + z;
    if (a < 0 || b < 0 || d < 0 || a >= g.dims.x || b >= g.dims.y || d >= g.dims.z) return kEmpty;
    return g.key_base + grid_linear_key(a, b, d, g);
}

// The reference hash is an XOR of per-axis products, and each axis term
// depends only on that coordinate and its stencil offset.  For the stencil
// (1,1,1) the three terms per axis are computed once (9 IEEE divisions
// instead of 81) and combined per cell; every key is the one hash_ref_xor
// gives for that cell.
struct XorAxisTerms {
    uint32_t x[3], y[3], z[3];
};
__device__ __forceinline__ uint32_t xor_axis_term(float p, float cs, float off, uint32_t prime) {
    return (uint32_t)f2i_sat(rn_div(rn_add(p + cs * off, 5000.0f), cs)) * prime;
}
__device__ __forceinline__ XorAxisTerms xor_axis_terms(const float4& sp, const GridDesc& g) {
    XorAxisTerms t;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float o = (float)(k - 1);
        t.x[k] = xor_axis_term(sp.x, g.cs.x, o, 73856093u);
        t.y[k] = xor_axis_term(sp.y, g.cs.y, o, 19349663u);
        t.z[k] = xor_axis_term(sp.z, g.cs.z, o, 83492791u);
    }
    return t;
}
__device__ __forceinline__ uint32_t sel3(const uint32_t (&v)[3], int k) {
    return k < 0 ? v[0] : (k > 0 ? v[2] : v[1]);
}
__device__ __forceinline__ uint32_t xor_key(const XorAxisTerms& t, int x, int y, int z, const GridDesc& g) {
    return g.key_base + ((sel3(t.x, x) ^ sel3(t.y, y) ^ sel3(t.z, z)) % g.nb);
}

// ProcessCollisionForcesD, K_DeviceFunctions.cu:113-159: the (2r+1)^3
// stencil in z, y, x order; REF_XOR grids rehash x + k*cs, LINEAR grids
// step cell coordinates and skip cells outside the grid.  For the usual
// relation (1,1,1) and PREFETCH the 27 cell starts are loaded together
// before any cell is walked (same cells, same order, same sums; more
// registers, so it is used by the body-sphere pass only).
template <bool PREFETCH, bool XOR_ONLY = false>
__device__ __forceinline__ F3 contact_stencil(const float4& sp, const float4& sv, float sr,
                                              const float4* __restrict__ cpos, const float4* __restrict__ cvel,
                                              float cr, float ks, float kd, int3 rel,
                                              const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                                              const GridDesc& g, bool w_index, bool group, int w_group, float eps,
                                              uint32_t limit, uint32_t vlimit) {
    F3 F = f3(0.0f, 0.0f, 0.0f);
    CellXYZ c{0, 0, 0};
    GridDesc gg = g;
    if (XOR_ONLY) gg.mode = FG_KEY_REF_XOR;  // body grids: only the hash is compiled in
    if (!XOR_ONLY && g.mode == FG_KEY_LINEAR) c = grid_cell(sp.x, sp.y, sp.z, g);
    const bool unit = rel.x == 1 && rel.y == 1 && rel.z == 1;
    const bool xor_terms = unit && gg.mode == FG_KEY_REF_XOR;
    XorAxisTerms t{};
    if (xor_terms) t = xor_axis_terms(sp, gg);
    if (PREFETCH && unit) {
        uint32_t keys[27], s0[27];
#pragma unroll
        for (int k = 0; k < 27; k++)
            keys[k] = xor_terms ? xor_key(t, k % 3 - 1, (k / 3) % 3 - 1, k / 9 - 1, gg)
                                : contact_key(sp, c, k % 3 - 1, (k / 3) % 3 - 1, k / 9 - 1, gg);
#pragma unroll
        for (int k = 0; k < 27; k++) s0[k] = keys[k] != kEmpty ? start[keys[k]] : kEmpty;
#pragma unroll
        for (int k = 0; k < 27; k++)
            if (keys[k] != kEmpty)  // an empty bucket adds +0, as in the reference
                F = add(F, contact_range(sp, sv, sr, cr, cpos, cvel, s0[k], s0[k] != kEmpty ? end[keys[k]] : 0u, ks,
                                         kd, w_index, group, w_group, eps, limit, vlimit));
        return F;
    }
    for (int z = -rel.z; z <= rel.z; z++)
        for (int y = -rel.y; y <= rel.y; y++)
            for (int x = -rel.x; x <= rel.x; x++) {
                const uint32_t key = xor_terms ? xor_key(t, x, y, z, gg) : contact_key(sp, c, x, y, z, gg);
                if (key == kEmpty) continue;
                const uint32_t s0 = start[key];
                F = add(F, contact_range(sp, sv, sr, cr, cpos, cvel, s0, s0 != kEmpty ? end[key] : 0u, ks, kd, w_index,
                                         group, w_group, eps, limit, vlimit));
            }
    return F;
}

// Force on a fluid particle from one body grid (K_FluidSim.cu:372-386);
// zero outside the expanded sphere bounds, where no contact is possible.
__device__ __forceinline__ F3 fluid_body_contact(const float4& pos, const float4& vel, const BodyContact& bc) {
    if (bc.aabb) {
        const float4 lo = bc.aabb[0], hi = bc.aabb[1];
        if (pos.x < lo.x || pos.y < lo.y || pos.z < lo.z || pos.x > hi.x || pos.y > hi.y || pos.z > hi.z)
            return f3(0.0f, 0.0f, 0.0f);
    }
    // body grids are keyed with the reference hash (body_grid in fg_rigid.hip)
    return contact_stencil<false, true>(pos, vel, bc.fluid_radius, bc.pos, bc.vel, bc.body_radius, bc.ks, bc.kd,
                                        bc.relation, bc.start, bc.end, bc.grid, bc.vel != nullptr, false, 0, bc.eps,
                                        bc.count, bc.nvel);
}

// contact_stencil (any relation, LINEAR or REF_XOR grid) with the cells of
// one sphere spread over LP consecutive lanes of a wavefront (LP a power of
// two <= 32; sub = the lane's index in its group; every lane of the group
// active and holding the same sphere).  Lane sub walks cells sub, sub + LP,
// ... of the z, y, x stencil order, each into a sum of its own from zero
// (contact_range), and after each round of LP cells the group adds the
// round's sums in cell order, as contact_stencil adds them.  Sums equal to
// zero are not added: the running sum starts at +0, so under
// round-to-nearest it is never -0, and adding +0 or -0 to it changes no
// bit.  The result is contact_stencil's, bit for bit, with the walks of
// different cells (chains of dependent loads: bucket range, then spheres)
// in flight together instead of one after another.
template <uint32_t LP>
__device__ __forceinline__ F3 contact_stencil_lanes(const float4& sp, const float4& sv, float sr,
                                                    const float4* __restrict__ cpos, const float4* __restrict__ cvel,
                                                    float cr, float ks, float kd, int3 rel,
                                                    const uint32_t* __restrict__ start,
                                                    const uint32_t* __restrict__ end, const GridDesc& g,
                                                    bool w_index, bool group, int w_group, float eps, uint32_t limit,
                                                    uint32_t vlimit, uint32_t sub) {
    static_assert(LP >= 1 && LP <= 32 && (LP & (LP - 1)) == 0, "LP: a power of two up to 32");
    F3 F = f3(0.0f, 0.0f, 0.0f);
    const CellXYZ c = g.mode == FG_KEY_LINEAR ? grid_cell(sp.x, sp.y, sp.z, g) : CellXYZ{0, 0, 0};
    const int nx = 2 * rel.x + 1, ny = 2 * rel.y + 1, cells = nx * ny * (2 * rel.z + 1);
    const bool unit = rel.x == 1 && rel.y == 1 && rel.z == 1;
    const uint32_t first = __lane_id() & ~(LP - 1u);
    for (int r0 = 0; r0 < cells; r0 += (int)LP) {
        const int k = r0 + (int)sub;
        F3 part = f3(0.0f, 0.0f, 0.0f);
        if (k < cells) {
            const int x = unit ? k % 3 - 1 : k % nx - rel.x;
            const int y = unit ? (k / 3) % 3 - 1 : (k / nx) % ny - rel.y;
            const int z = unit ? k / 9 - 1 : k / (nx * ny) - rel.z;
            const uint32_t key = contact_key(sp, c, x, y, z, g);
            if (key != kEmpty) {  // a LINEAR cell outside the grid adds nothing
                const uint32_t s0 = start[key], e = end[key];
                part = contact_range(sp, sv, sr, cr, cpos, cvel, s0, s0 != kEmpty ? e : 0u, ks, kd, w_index, group,
                                     w_group, eps, limit, vlimit);
            }
        }
        const uint64_t nz = __ballot(part.x != 0.0f || part.y != 0.0f || part.z != 0.0f);
        uint32_t m = (uint32_t)(nz >> first) & (LP == 32 ? 0xFFFFFFFFu : ((1u << LP) - 1u));
        while (m) {
            const int src = (int)first + __ffs((int)m) - 1;
            m &= m - 1u;
            F = add(F, f3(__shfl(part.x, src), __shfl(part.y, src), __shfl(part.z, src)));
        }
    }
    return F;
}

// fluid_body_contact over LP lanes (contact_stencil_lanes).
template <uint32_t LP>
__device__ __forceinline__ F3 contact_cells(const float4& pos, const float4& vel, const BodyContact& bc,
                                            uint32_t sub) {
    if (bc.aabb) {
        const float4 lo = bc.aabb[0], hi = bc.aabb[1];
        if (pos.x < lo.x || pos.y < lo.y || pos.z < lo.z || pos.x > hi.x || pos.y > hi.y || pos.z > hi.z)
            return f3(0.0f, 0.0f, 0.0f);
    }
    GridDesc g = bc.grid;
    g.mode = FG_KEY_REF_XOR;  // body grids are keyed with the reference hash (body_grid in fg_rigid.hip)
    return contact_stencil_lanes<LP>(pos, vel, bc.fluid_radius, bc.pos, bc.vel, bc.body_radius, bc.ks, bc.kd,
                                     bc.relation, bc.start, bc.end, g, bc.vel != nullptr, false, 0, bc.eps, bc.count,
                                     bc.nvel, sub);
}

// External force of K_FluidSim.cu:369-386: static grid first, then the
// dynamic grid, summed from zero.
__device__ __forceinline__ F3 fluid_external(const float4& pos, const float4& vel, const Contacts& c) {
    F3 e = f3(0.0f, 0.0f, 0.0f);
    if (c.st.enabled) e = add(e, fluid_body_contact(pos, vel, c.st));
    if (c.dyn.enabled) e = add(e, fluid_body_contact(pos, vel, c.dyn));
    return e;
}

#pragma clang fp contract(fast)
}  // namespace fg
