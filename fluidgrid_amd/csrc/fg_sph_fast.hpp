// fg_sph_fast.hpp -- per-particle SPH work shared by the single-fluid
// kernels (fg_sph.hip) and the batched-scene kernels (fg_batch.hip):
//  * acceleration after gravity, walls, body contact and clamp, and the
//    velocity/position write-back (reference arithmetic, contraction off);
//  * the fast LINEAR density and force passes over the 9 stencil rows.
#pragma once
#include "fg_contact.hpp"
#include "fg_internal.hpp"
#include "fg_walls.hpp"

namespace fg {

#pragma clang fp contract(off)

// Walls and clamp of K_FluidSim.cu:293-303 applied to the acceleration
// `total` (pair sum times mass, gravity and any body contact).
__device__ __forceinline__ F3 walls_and_clamp(F3 total, const float4& pos, const float4& vel,
                                              const fg_common_fluid_params& cp) {
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

// Acceleration after gravity, walls and clamp (K_FluidSim.cu:287-303), from
// the pair sum `internal`.  NO_CONTACT: the caller knows the particle lies
// outside every body grid's bounds, where fluid_external is exactly +0.
template <bool COUPLED, bool NO_CONTACT = false>
__device__ __forceinline__ F3 finish_accel(F3 internal, float m, const float4& pos, const float4& vel,
                                           const fg_common_fluid_params& cp, const Contacts& bc) {
    const F3 g = f3(cp.global_external_force.x, cp.global_external_force.y, cp.global_external_force.z);
    F3 total = add(mul(internal, m), g);
    // K_FluidSim.cu:389-390: body contact forces are added after the mass
    // factor (the uncoupled kernel adds nothing here).
    if (COUPLED && bc.coupled) total = add(total, NO_CONTACT ? f3(0.0f, 0.0f, 0.0f) : fluid_external(pos, vel, bc));
    return walls_and_clamp(total, pos, vel, cp);
}

// The same with the contact term `ext` computed by the caller (added when
// `coupled`, as fluid_external is above).
__device__ __forceinline__ F3 finish_accel_ext(F3 internal, float m, const float4& pos, const float4& vel,
                                               const fg_common_fluid_params& cp, bool coupled, F3 ext) {
    const F3 g = f3(cp.global_external_force.x, cp.global_external_force.y, cp.global_external_force.z);
    F3 total = add(mul(internal, m), g);
    if (coupled) total = add(total, ext);
    return walls_and_clamp(total, pos, vel, cp);
}

// Whether body contact can act on a particle at pos: it lies inside the
// expanded bounds of an enabled body grid (or the grid has none), the test
// with which fluid_body_contact skips its walk.
__device__ __forceinline__ bool contact_possible(const float4& pos, const Contacts& c) {
    auto inside = [&](const BodyContact& b) -> bool {
        if (!b.enabled) return false;
        if (!b.aabb) return true;
        const float4 lo = b.aabb[0], hi = b.aabb[1];
        return !(pos.x < lo.x || pos.y < lo.y || pos.z < lo.z || pos.x > hi.x || pos.y > hi.y || pos.z > hi.z);
    };
    return inside(c.st) || inside(c.dyn);
}

__device__ __forceinline__ void defer_particle(const DeferredContact& d, uint32_t i, F3 internal) {
    d.acc[i] = make_float4(internal.x, internal.y, internal.z, 0.0f);
    // one atomic per wavefront: the deferring lanes take consecutive slots
    const uint64_t lanes = __ballot(1);
    const uint32_t leader = (uint32_t)__ffsll((long long)lanes) - 1u;
    uint32_t base = 0;
    if (__lane_id() == leader) base = atomicAdd(d.count, (uint32_t)__popcll(lanes));
    base = (uint32_t)__shfl((int)base, (int)leader);
    const uint32_t lane = __lane_id();
    const uint64_t below = lane == 0 ? 0ull : (lanes & ((~0ull) >> (64 - lane)));
    d.list[base + (uint32_t)__popcll(below)] = i;
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

#pragma clang fp contract(fast)

// ---------------------------------------------------------------------------
// Fast LINEAR kernels
// ---------------------------------------------------------------------------

// Per-particle view of the 3x3 stencil rows.  A row is the x-run of cells
// cx-1..cx+1 at (cy+dy, cz+dz); its particles are sorted[row.begin, row.end).
struct RowWalker {
    float gx_lo, gx_hi;  // squared scaled distance to the x-neighbour cells
    float gy_lo, gy_hi;  // ... to the y-neighbour slabs (0 for the own slab)
    float gz_lo, gz_hi;  // ... to the z-neighbour slabs
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
    w.gy_lo = sq(fmaxf(0.0f, p.y - ly) * s);
    w.gy_hi = sq(fmaxf(0.0f, hy - p.y) * s);
    w.gz_lo = sq(fmaxf(0.0f, p.z - lz) * s);
    w.gz_hi = sq(fmaxf(0.0f, hz - p.z) * s);
    return w;
}

// Sorted range of row (dy,dz) after culling, in two halves so that a row's
// loads can be issued a row ahead (row_fetch) and used later (row_finish).
// The grid's buckets start at key_base of the key space (0 for a single
// fluid; the scene's first bucket in a batch).  `cut` is h^2 enlarged by a
// small margin so that rounding in the cell bounds can never drop a pair
// with r < h.
struct RowFetch {
    uint32_t s0, s1, e1;  // start of the first cell, start and end of the last
    bool ok;              // the row is not culled (and s0, s1, e1 are loaded)
};

// The row's cells cx-1..cx+1 after culling by distance, as keys k0..k1;
// false for a culled row.
__device__ __forceinline__ bool row_keys(const RowWalker& w, int dy, int dz, float cut,
                                         const fg_common_fluid_params& cp, uint32_t key_base, uint32_t& k0,
                                         uint32_t& k1) {
    const int cy = w.cy + dy, cz = w.cz + dz;
    if (cy < 0 || cz < 0 || cy >= cp.bucket_size.y || cz >= cp.bucket_size.z) return false;
    // selects, not an array indexed by the row: the row loops are not always
    // unrolled, and an indexed array would live in scratch memory
    const float gy = dy < 0 ? w.gy_lo : (dy > 0 ? w.gy_hi : 0.0f);
    const float gz = dz < 0 ? w.gz_lo : (dz > 0 ? w.gz_hi : 0.0f);
    const float gyz = gy + gz;
    if (gyz >= cut) return false;
    const bool useL = w.cx > 0 && gyz + w.gx_lo < cut;                      // cell cx-1
    const bool useR = w.cx + 1 < cp.bucket_size.x && gyz + w.gx_hi < cut;  // cell cx+1
    const uint32_t kb = key_base + linear_key(0, cy, cz, cp) + (uint32_t)w.cx;
    k0 = useL ? kb - 1u : kb;
    k1 = useR ? kb + 1u : kb;
    return true;
}

// The row's particles are sorted[start[k0], end[k1]) when its end cells are
// occupied, as in the bulk of the fluid: three loads in one round trip.
__device__ __forceinline__ RowFetch row_fetch(const RowWalker& w, int dy, int dz, float cut,
                                              const fg_common_fluid_params& cp, uint32_t key_base,
                                              const uint32_t* __restrict__ start, const uint32_t* __restrict__ end) {
    RowFetch f{kEmpty, kEmpty, 0u, false};
    uint32_t k0, k1;
    if (row_keys(w, dy, dz, cut, cp, key_base, k0, k1)) {
        f.s0 = start[k0];
        f.s1 = start[k1];
        f.e1 = end[k1];
        f.ok = true;
    }
    return f;
}

// The range from the fetched loads; when an end cell is empty, the first
// occupied cell from the left and the last from the right bound it (end[]
// of an empty cell is stale and unused).  Returns false when empty.
__device__ __forceinline__ bool row_finish(const RowFetch& f, const RowWalker& w, int dy, int dz, float cut,
                                           const fg_common_fluid_params& cp, uint32_t key_base,
                                           const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                                           uint32_t n, uint32_t& rb, uint32_t& re) {
    if (!f.ok) return false;
    if (f.s0 != kEmpty && f.s1 != kEmpty) {
        rb = f.s0;
        re = min(f.e1, n);
        return rb < re;
    }
    uint32_t k0 = 0, k1 = 0;
    row_keys(w, dy, dz, cut, cp, key_base, k0, k1);
    uint32_t b = kEmpty, e = 0;
    for (uint32_t k = k0; k <= k1; k++) {
        const uint32_t sk = start[k];
        if (sk != kEmpty) {
            if (b == kEmpty) b = sk;
            e = end[k];
        }
    }
    if (b == kEmpty) return false;
    rb = b;
    re = min(e, n);
    return rb < re;
}

__device__ __forceinline__ bool row_range(const RowWalker& w, int dy, int dz, float cut,
                                          const fg_common_fluid_params& cp, uint32_t key_base,
                                          const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                                          uint32_t n, uint32_t& rb, uint32_t& re) {
    return row_finish(row_fetch(w, dy, dz, cut, cp, key_base, start, end), w, dy, dz, cut, cp, key_base, start,
                      end, n, rb, re);
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

// After the 9 row masks, at the next 256-byte boundary, the neighbour records
// the force pass reads for every neighbour j: two float4 per particle,
// (x_j, y_j, z_j, 1/rho_j) and (vx_j, vy_j, vz_j, p_j/rho_j), all from the
// sorted arrays.  One 32-byte record takes two 16-byte loads from one cache
// line, where position, velocity and (1/rho, p/rho) read from three arrays
// took three loads touching three lines.
__host__ __device__ __forceinline__ size_t nbr_record_offset(uint32_t n) {
    return ((size_t)9 * sizeof(unsigned long long) * n + 255) & ~(size_t)255;
}
__device__ __forceinline__ float4* nbr_records(unsigned long long* masks, uint32_t n) {
    return reinterpret_cast<float4*>(reinterpret_cast<char*>(masks) + nbr_record_offset(n));
}
__device__ __forceinline__ const float4* nbr_records(const unsigned long long* masks, uint32_t n) {
    return reinterpret_cast<const float4*>(reinterpret_cast<const char*>(masks) + nbr_record_offset(n));
}

// xyz of a float4 array element as one 12-byte load (the w of a neighbour
// is not used by the pair terms).
__device__ __forceinline__ float3 ld_xyz(const float4* p) {
    const float* f = reinterpret_cast<const float*>(p);
    return make_float3(f[0], f[1], f[2]);
}

// LDS staging of neighbour rows, per wavefront.  The lanes of a wavefront
// are consecutive sorted particles, so for each of the 9 stencil rows the
// union of their candidate ranges is one short run of sorted indices
// [lo, hi): the wavefront loads it once, coalesced, into its own LDS slice
// and the lanes read their candidates from there instead of gathering them
// from memory, where the lanes of a row touch the same particles many
// times.  A row whose run does not fit its slice is read from memory.  The
// density pass stages positions (kStagePos float4 per wavefront), the force
// pass neighbour records (kStageRecs records of two float4).
constexpr uint32_t kStagePos = 256;
constexpr uint32_t kStageRecs = 128;

// Orders a wavefront's LDS stores before its other lanes' loads (and loads
// before the next stores): the lanes of one wavefront need no workgroup
// barrier.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Whether one row is staged for the wavefront, and its run [lo, hi).  The
// sorted ranges of consecutive lanes are ordered (a later lane's cell has a
// key no lower), so the run is the start of the first lane with the row and
// the end of the last one: two lane reads, where a min/max reduction over 64
// lanes is six dependent cross-lane steps.  A lane whose range falls outside
// that run sends the whole row to memory, so the result never depends on
// the ordering.  Every lane of the wavefront must call it.
__device__ __forceinline__ bool wave_run(bool has, uint32_t rb, uint32_t re, uint32_t cap, uint32_t& lo,
                                         uint32_t& hi) {
    const unsigned long long hm = __ballot(has);
    lo = hi = 0u;
    if (hm == 0ull) return false;
    lo = (uint32_t)__builtin_amdgcn_readlane((int)rb, __builtin_ctzll(hm));
    hi = (uint32_t)__builtin_amdgcn_readlane((int)re, 63 - __builtin_clzll(hm));
    const bool inside = !has || (rb >= lo && re <= hi);
    return __ballot(!inside) == 0ull && hi - lo <= cap;
}

// The same run as a min/max reduction over the 64 lanes (six dependent
// cross-lane steps, no lane order assumed).  The force pass uses it: with
// the two lane reads above it measured 7% slower (0.88 against 0.82 ms at
// 8,388,608 particles, same bits), and at 8 waves per SIMD they also cost a
// register spill reloaded at every row.
__device__ __forceinline__ bool wave_run_minmax(bool has, uint32_t rb, uint32_t re, uint32_t cap, uint32_t& lo,
                                                uint32_t& hi) {
    lo = has ? rb : 0xFFFFFFFFu;
    hi = has ? re : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, o));
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, o));
    }
    return hi > lo && hi - lo <= cap;
}

// Poly6 sum over four candidates; bit u of the returned nibble is set when
// candidate u lies within h.
//
// The fused multiply-adds are written out (and contraction is off here), so
// every schedule and workgroup size that inlines this evaluates exactly the
// same operations.
#pragma clang fp contract(off)
__device__ __forceinline__ uint32_t density4(const float4& p, const float3& q0, const float3& q1, const float3& q2,
                                             const float3& q3, float h2u, float hhu, float& acc) {
    const float3 q[4] = {q0, q1, q2, q3};
    uint32_t nib = 0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const float dx = p.x - q[u].x, dy = p.y - q[u].y, dz = p.z - q[u].z;
        const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
        const float t = h2u - r2;
        const bool hit = r2 < hhu;
        acc = hit ? __builtin_fmaf(t * t, t, acc) : acc;
        nib |= hit ? (1u << u) : 0u;
    }
    return nib;
}
#pragma clang fp contract(fast)

// rho_p[i] = (max(1, rho), k (rho - rho0)) (K_FluidSim.cu:246-249) and, when
// rec is given, particle i's neighbour record from its sorted position p,
// sorted velocity v and (1/rho_p.x, rho_p.y/rho_p.x).
#pragma clang fp contract(off)
__device__ __forceinline__ void store_record(uint32_t i, float2 rp, const float4& p, const float4& v,
                                             float4* __restrict__ rec) {
    const float a = 1.0f / rp.x;
    rec[2 * (size_t)i] = make_float4(p.x, p.y, p.z, a);
    rec[2 * (size_t)i + 1] = make_float4(v.x, v.y, v.z, rp.y * a);
}
__device__ __forceinline__ void store_density(uint32_t i, float rho, float k, float rho0, float2* __restrict__ rho_p,
                                              float4* __restrict__ rec, const float4& p, const float4& v) {
    const float2 rp = make_float2(fmaxf(1.0f, rho), k * (rho - rho0));
    rho_p[i] = rp;
    if (rec) store_record(i, rp, p, v, rec);
}
#pragma clang fp contract(fast)

// xyz of a position staged in LDS (a 16-byte read; w unused).
__device__ __forceinline__ float3 ld_xyz_lds(const float4* p) {
    const float4 q = *p;
    return make_float3(q.x, q.y, q.z);
}

// One stencil row of the density pass: the len candidates src[0, len) in
// groups of four in sorted order (a short last group padded with far
// points) added to acc, and the row's mask (0 for rows longer than 64
// candidates, which are never read through their mask).  src is the sorted
// positions from the row's first particle, in memory or (LDS) staged.
template <bool MASKS, bool LDS>
__device__ __forceinline__ unsigned long long density_row(const float4& p, const float4* __restrict__ pp,
                                                          uint32_t len, float h2u, float hhu, float& acc) {
    auto ld = [](const float4* q) { return LDS ? ld_xyz_lds(q) : ld_xyz(q); };
    unsigned long long msk = 0ull;
    const uint32_t n4 = len & ~3u;
    uint32_t kk = 0;
    if (LDS) {
        // From LDS the loads return quickly: no pipelining, whose register
        // copies cost a quarter of the loop's instructions.
        for (; kk < n4; kk += 4, pp += 4) {
            const uint32_t nib = density4(p, ld(pp), ld(pp + 1), ld(pp + 2), ld(pp + 3), h2u, hhu, acc);
            if (MASKS) msk |= (unsigned long long)nib << (kk & 63u);
        }
    } else if (n4) {
        // Software-pipelined groups of four: the next group's loads are
        // in flight while the current group is evaluated.
        float3 a0 = ld(pp), a1 = ld(pp + 1), a2 = ld(pp + 2), a3 = ld(pp + 3);
        for (; kk + 4 < n4; kk += 4) {
            pp += 4;
            const float3 b0 = ld(pp), b1 = ld(pp + 1), b2 = ld(pp + 2), b3 = ld(pp + 3);
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
        const float3 far = make_float3(3.0e38f, 3.0e38f, 3.0e38f);
        const float3 t0 = ld(pp);
        const float3 t1 = rest > 1 ? ld(pp + 1) : far;
        const float3 t2 = rest > 2 ? ld(pp + 2) : far;
        const uint32_t nib = density4(p, t0, t1, t2, far, h2u, hhu, acc);
        if (MASKS) msk |= (unsigned long long)nib << (kk & 63u);
    }
    return len <= kRowMaskBits ? msk : 0ull;
}

// The row loop of the density pass with each row's candidate positions
// staged in the wavefront's LDS slice `stage` (kStagePos positions), for a
// whole wavefront (every lane present; a lane of an inactive particle takes
// part with no rows and writes no masks).  Same candidates, same order,
// same sums and masks as the row loop of density_particle.
template <bool MASKS>
__device__ __forceinline__ float density_rows_staged(uint32_t i, bool active, const float4& p,
                                                     const float4* __restrict__ sp,
                                                     const uint32_t* __restrict__ start,
                                                     const uint32_t* __restrict__ end, uint32_t n,
                                                     const fg_common_fluid_params& cp, uint32_t key_base,
                                                     unsigned long long* __restrict__ masks,
                                                     float4* __restrict__ stage) {
    const float s = cp.simulation_scale, s2 = s * s;
    const float h = cp.kernel_radius, hh = h * h;
    const float hu = h / s, hhu = hu * hu, h2u = cp.kernel_radius_squared / s2;
    const float cut = hh * 1.0001f;
    const RowWalker w = make_rows(p, cp);
    float acc = 0.0f;
    // The next row's cell bounds are loaded a row ahead, after this row's
    // staging loads, so they arrive while this row is walked.
    RowFetch next = active ? row_fetch(w, -1, -1, cut, cp, key_base, start, end) : RowFetch{kEmpty, kEmpty, 0u, false};
    for (int row = 0; row < 9; row++) {
        const int dz = row / 3 - 1, dy = row % 3 - 1;
        const RowFetch f = next;
        uint32_t rb = 0, re = 0;
        const bool has = active && row_finish(f, w, dy, dz, cut, cp, key_base, start, end, n, rb, re);
        uint32_t lo, hi;
        const bool staged = wave_run(has, rb, re, kStagePos, lo, hi);  // the same on every lane
        if (staged)
            for (uint32_t k = __lane_id(); k < hi - lo; k += 64u) stage[k] = sp[lo + k];
        if (row < 8 && active) next = row_fetch(w, (row + 1) % 3 - 1, (row + 1) / 3 - 1, cut, cp, key_base, start, end);
        if (staged) wave_lds_sync();
        unsigned long long msk = 0ull;
        if (has)
            msk = staged ? density_row<MASKS, true>(p, stage + (rb - lo), re - rb, h2u, hhu, acc)
                         : density_row<MASKS, false>(p, sp + rb, re - rb, h2u, hhu, acc);
        if (MASKS && active) masks[(size_t)row * n + i] = msk;
        if (staged) wave_lds_sync();  // every lane's reads of this row before the next row's stores
    }
    return acc;
}

// Per-lane row table of the flat density schedule, in LDS: first and
// one-past-last sorted index of each of the 9 rows (equal for a culled row).
struct FlatRanges {
    uint32_t* rb;  // [9][lanes]
    uint32_t* re;  // [9][lanes]
    uint32_t lanes, lane;
    __device__ __forceinline__ uint32_t& b(int r) { return rb[r * lanes + lane]; }
    __device__ __forceinline__ uint32_t& e(int r) { return re[r * lanes + lane]; }
};

// Density of sorted particle i (i < n): rho_p[i] = (max(1, rho), k (rho - rho0)),
// and, with MASKS, its row masks and its neighbour record (sv, the sorted
// velocities, is read only then).
//
// Two schedules give bit-identical sums (rows in order, candidates of a row
// in groups of four in sorted order, a short last group padded with far
// points): the row loop, where a wavefront waits at every row for its lane
// with the longest row, and the flat schedule (with a row table), where each
// lane walks its own candidate stream across rows.  Given an LDS slice
// (`stage`) and a full wavefront, the row loop reads each row's candidates
// from LDS (density_rows_staged).
template <bool MASKS>
__device__ __forceinline__ void density_particle(uint32_t i, const float4* __restrict__ sp,
                                                 const float4* __restrict__ sv, float2* __restrict__ rho_p, const uint32_t* __restrict__ start,
                                                 const uint32_t* __restrict__ end, uint32_t n, float k, float rho0,
                                                 float m, const fg_common_fluid_params& cp,
                                                 const fg_kernels_cache& kc, uint32_t key_base,
                                                 unsigned long long* __restrict__ masks,
                                                 FlatRanges* flat = nullptr, float4* __restrict__ stage = nullptr) {
    const float4 p = sp[i];
    float4* rec = MASKS ? nbr_records(masks, n) : nullptr;
    const float4 v = MASKS ? sv[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // The staged row loop needs every lane of the wavefront (the run of each
    // row is a wavefront reduction); otherwise the rows are read from memory.
    if (stage != nullptr && flat == nullptr && __ballot(1) == ~0ull) {
        const bool active = p.w > 0.0f;
        const float acc = density_rows_staged<MASKS>(i, active, p, sp, start, end, n, cp, key_base, masks, stage);
        if (active) {
            const float s2 = cp.simulation_scale * cp.simulation_scale;
            store_density(i, m * kc.kPoly6 * (s2 * s2 * s2) * acc, k, rho0, rho_p, rec, p, v);
        } else if (MASKS) {
            store_record(i, rho_p[i], p, v, rec);
        }
        return;
    }
    if (!(p.w > 0.0f)) {
        // (rho, p) is left as it is; neighbours still read it
        if (MASKS) store_record(i, rho_p[i], p, v, rec);
        return;
    }
    // Distances are compared unscaled: r_u = r / s, so r < h <=> r_u^2 < (h/s)^2
    // and h^2 - r^2 = s^2 (h^2/s^2 - r_u^2).
    const float s = cp.simulation_scale, s2 = s * s;
    const float h = cp.kernel_radius, hh = h * h;
    const float hu = h / s, hhu = hu * hu, h2u = cp.kernel_radius_squared / s2;
    const float cut = hh * 1.0001f;
    const RowWalker w = make_rows(p, cp);
    float acc = 0.0f;
    if (flat) {
#pragma unroll
        for (int row = 0; row < 9; row++) {
            const int dz = row / 3 - 1, dy = row % 3 - 1;
            uint32_t rb = 0, re = 0;
            if (!row_range(w, dy, dz, cut, cp, key_base, start, end, n, rb, re)) rb = re = 0;
            flat->b(row) = rb;
            flat->e(row) = re;
        }
        const float3 far = make_float3(3.0e38f, 3.0e38f, 3.0e38f);
        int row = -1;
        uint32_t j = 0, rb = 0, re = 0;
        unsigned long long msk = 0ull;
        for (;;) {
            while (j >= re) {
                if (MASKS && row >= 0 && re > rb) masks[(size_t)row * n + i] = msk;
                if (row == 8) break;
                row++;
                rb = flat->b(row);
                re = flat->e(row);
                j = rb;
                msk = 0ull;
            }
            if (j >= re) break;
            const uint32_t rest = re - j;
            const float4* pp = sp + j;
            const float3 t0 = ld_xyz(pp);
            const float3 t1 = rest > 1 ? ld_xyz(pp + 1) : far;
            const float3 t2 = rest > 2 ? ld_xyz(pp + 2) : far;
            const float3 t3 = rest > 3 ? ld_xyz(pp + 3) : far;
            const uint32_t nib = density4(p, t0, t1, t2, t3, h2u, hhu, acc);
            if (MASKS) msk |= (unsigned long long)nib << ((j - rb) & 63u);
            j += 4;
        }
        store_density(i, m * kc.kPoly6 * (s2 * s2 * s2) * acc, k, rho0, rho_p, rec, p, v);
        return;
    }
#pragma unroll
    for (int row = 0; row < 9; row++) {
        const int dz = row / 3 - 1, dy = row % 3 - 1;
        unsigned long long msk = 0ull;
        uint32_t rb, re;
        if (row_range(w, dy, dz, cut, cp, key_base, start, end, n, rb, re))
            msk = density_row<MASKS, false>(p, sp + rb, re - rb, h2u, hhu, acc);
        if (MASKS) masks[(size_t)row * n + i] = msk;
    }
    store_density(i, m * kc.kPoly6 * (s2 * s2 * s2) * acc, k, rho0, rho_p, rec, p, v);
}

// Pressure and viscosity terms of one neighbour j (r < h checked again, so
// the same code serves rows without a mask), from aj = 1/rho_j and
// bj = p_j/rho_j: (p_i + p_j)/rho_j = p_i aj + bj.
struct PairAcc {
    float px, py, pz, vx, vy, vz;
};

#pragma clang fp contract(off)
// The accumulation of one pair with r2u = |x_i - x_j|^2 (unscaled) < h^2.
__device__ __forceinline__ void pair_accumulate(PairAcc& A, float dx, float dy, float dz, float r2u,
                                                const float4& vel, float pri, float3 vj, float aj, float bj,
                                                float h, float s) {
    const float iru = __builtin_amdgcn_rsqf(r2u);
    const float hr = __builtin_fmaf(-(r2u * iru), s, h);
    const float cp_ = hr * hr * __builtin_fmaf(pri, aj, bj) * iru;
    A.px = __builtin_fmaf(dx, cp_, A.px);
    A.py = __builtin_fmaf(dy, cp_, A.py);
    A.pz = __builtin_fmaf(dz, cp_, A.pz);
    const float cv = hr * aj;
    A.vx = __builtin_fmaf(vj.x - vel.x, cv, A.vx);
    A.vy = __builtin_fmaf(vj.y - vel.y, cv, A.vy);
    A.vz = __builtin_fmaf(vj.z - vel.z, cv, A.vz);
}

__device__ __forceinline__ void pair_term(PairAcc& A, uint32_t i, uint32_t j, const float4& pos, const float4& vel,
                                          float pri, float3 q, float3 vj, float aj, float bj, float h, float s,
                                          float hhu) {
    const float dx = pos.x - q.x, dy = pos.y - q.y, dz = pos.z - q.z;
    const float r2u = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));  // unscaled
    if (r2u < hhu && j != i) pair_accumulate(A, dx, dy, dz, r2u, vel, pri, vj, aj, bj, h, s);
}

// A pair from a neighbour mask (self bit cleared).  The density pass set the
// bit after the same test on the same r2u (the same operations on the same
// sorted positions), so r2u < h^2 holds and is not tested again: without
// the branch the neighbour's velocity and inverse-density loads are issued
// together with its position load instead of after the test.
__device__ __forceinline__ void pair_term_masked(PairAcc& A, const float4& pos, const float4& vel, float pri,
                                                 float3 q, float3 vj, float aj, float bj, float h, float s) {
    const float dx = pos.x - q.x, dy = pos.y - q.y, dz = pos.z - q.z;
    const float r2u = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    pair_accumulate(A, dx, dy, dz, r2u, vel, pri, vj, aj, bj, h, s);
}

// The same from a stored (rho_j, p_j) (rows without a mask), with the
// (1/rho_j, p_j/rho_j) the density pass stores beside the masks.
__device__ __forceinline__ void pair_term_rp(PairAcc& A, uint32_t i, uint32_t j, const float4& pos,
                                             const float4& vel, float pri, float3 q, float3 vj, float2 dpj, float h,
                                             float s, float hhu) {
    const float aj = 1.0f / dpj.x;
    pair_term(A, i, j, pos, vel, pri, q, vj, aj, dpj.y * aj, h, s, hhu);
}
#pragma clang fp contract(fast)

// The neighbours of one row from its mask mm (self bit cleared), bit k
// being record first + k, in ascending order, two per iteration.  Record j
// is (pa[STRIDE j], pb[STRIDE j]): the interleaved records in memory
// (STRIDE 2, pb = pa + 1) or the two halves staged apart in LDS (STRIDE 1),
// where 16-byte reads of consecutive records then fall in distinct banks.
// The mask is walked as two 32-bit words, low word first (the same
// ascending order): 32-bit bit scans and clears, and rows of at most 32
// candidates never enter the second loop.
template <int STRIDE>
__device__ __forceinline__ void walk_row_mask(PairAcc& A, unsigned long long mm, uint32_t first,
                                              const float4* __restrict__ pa, const float4* __restrict__ pb,
                                              const float4& pos, const float4& vel, float pri, float h, float s) {
    uint32_t word = (uint32_t)mm, base = first;
    for (int half = 0; half < 2; half++) {
        while (word) {
            const uint32_t k1 = (uint32_t)__builtin_ctz(word);
            word &= word - 1u;
            const bool two = word != 0u;
            const uint32_t k2 = two ? (uint32_t)__builtin_ctz(word) : k1;
            word &= word - 1u;  // (0 stays 0)
            const uint32_t j1 = base + k1, j2 = base + k2;
            const float4 a1 = pa[STRIDE * (size_t)j1], b1 = pb[STRIDE * (size_t)j1];
            const float4 a2 = pa[STRIDE * (size_t)j2], b2 = pb[STRIDE * (size_t)j2];
            const float3 q1 = make_float3(a1.x, a1.y, a1.z), q2 = make_float3(a2.x, a2.y, a2.z);
            const float3 v1 = make_float3(b1.x, b1.y, b1.z), v2 = make_float3(b2.x, b2.y, b2.z);
            const float2 d1 = make_float2(a1.w, b1.w), d2 = make_float2(a2.w, b2.w);
            // Keeps the second pair's loads beside the first pair's (one
            // memory round trip per iteration): without a use here the
            // compiler sinks them into the `two` branch below.
            asm volatile("" ::"v"(q2.x), "v"(q2.y), "v"(q2.z), "v"(v2.x), "v"(v2.y), "v"(v2.z), "v"(d2.x),
                         "v"(d2.y));
            pair_term_masked(A, pos, vel, pri, q1, v1, d1.x, d1.y, h, s);
            if (two) pair_term_masked(A, pos, vel, pri, q2, v2, d2.x, d2.y, h, s);
        }
        word = (uint32_t)(mm >> 32);  // then the high word
        base = first + 32u;
        if (!word) break;
    }
}

// The row loop of the force pass: the 9 stencil rows in turn, each row's
// neighbours from its mask (ascending, two per iteration) or, without a mask
// or for a row longer than 64 candidates, by re-testing every candidate.
template <bool MASKS>
__device__ __forceinline__ void forces_rowloop(PairAcc& A, uint32_t i, const float4& pos, const float4& vel,
                                               float2 dpi, const float4* __restrict__ sp,
                                               const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                                               const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                                               uint32_t n, const fg_common_fluid_params& cp, uint32_t key_base,
                                               const unsigned long long* __restrict__ masks) {
    const float h = cp.kernel_radius, hh = h * h;
    const float s = cp.simulation_scale;
    const float hu = h / s, hhu = hu * hu;
    const float cut = hh * 1.0001f;
    const RowWalker w = make_rows(pos, cp);
    const float4* rec = MASKS ? nbr_records(masks, n) : nullptr;
    for (int row = 0; row < 9; row++) {
        const int dz = row / 3 - 1, dy = row % 3 - 1;
        // The row's mask is read before its range is known, so the two
        // reads share one memory round trip (the asm keeps the compiler
        // from sinking the load behind the range test).
        unsigned long long mrow = 0ull;
        if (MASKS) {
            mrow = masks[(size_t)row * n + i];
            asm volatile("" ::"v"(mrow));
        }
        uint32_t rb, re;
        if (!row_range(w, dy, dz, cut, cp, key_base, start, end, n, rb, re)) continue;
        if (MASKS && re - rb <= kRowMaskBits) {
            unsigned long long mm = mrow;
            // the particle itself is in its own row's mask; it adds nothing
            if (i >= rb && i < re) mm &= ~(1ull << (i - rb));
            walk_row_mask<2>(A, mm, rb, rec, rec + 1, pos, vel, dpi.y, h, s);
        } else {
            for (uint32_t j = rb; j < re; j++)
                pair_term_rp(A, i, j, pos, vel, dpi.y, ld_xyz(sp + j), ld_xyz(sv + j), rho_p[j], h, s, hhu);
        }
    }
}

// The masked row loop with each row's neighbour records staged in the
// wavefront's LDS slice `stage` (kStageRecs records: their first halves,
// then their second halves), for a whole wavefront
// (every lane present; a lane of an inactive particle takes part with no
// rows).  Same pairs, same order, same sums as forces_rowloop<true>.
// AHEAD: the next row's cell bounds and mask are loaded a row ahead, after
// this row's staging loads, so they arrive while this row is walked (about
// five more registers; the single-fluid kernel runs it at 7 waves per SIMD).
template <bool AHEAD>
__device__ __forceinline__ void forces_rows_staged(PairAcc& A, uint32_t i, bool active, const float4& pos,
                                                   const float4& vel, float2 dpi, const float4* __restrict__ sp,
                                                   const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                                                   const uint32_t* __restrict__ start,
                                                   const uint32_t* __restrict__ end, uint32_t n,
                                                   const fg_common_fluid_params& cp, uint32_t key_base,
                                                   const unsigned long long* __restrict__ masks,
                                                   float4* __restrict__ stage) {
    const float h = cp.kernel_radius, hh = h * h;
    const float s = cp.simulation_scale;
    const float hu = h / s, hhu = hu * hu;
    const float cut = hh * 1.0001f;
    const RowWalker w = make_rows(pos, cp);
    const float4* rec = nbr_records(masks, n);
    RowFetch next{kEmpty, kEmpty, 0u, false};
    unsigned long long next_mask = 0ull;
    if (AHEAD && active) {
        next = row_fetch(w, -1, -1, cut, cp, key_base, start, end);
        next_mask = masks[i];
    }
    for (int row = 0; row < 9; row++) {
        const int dz = row / 3 - 1, dy = row % 3 - 1;
        unsigned long long mrow = 0ull;
        uint32_t rb = 0, re = 0;
        bool has = false;
        if (AHEAD) {
            mrow = next_mask;
            has = active && row_finish(next, w, dy, dz, cut, cp, key_base, start, end, n, rb, re);
        } else if (active) {
            mrow = masks[(size_t)row * n + i];
            asm volatile("" ::"v"(mrow));
            has = row_range(w, dy, dz, cut, cp, key_base, start, end, n, rb, re);
        }
        uint32_t lo, hi;
        const bool staged = wave_run_minmax(has, rb, re, kStageRecs, lo, hi);  // the same on every lane
        if (staged) {
            for (uint32_t k = __lane_id(); k < hi - lo; k += 64u) {
                stage[k] = rec[2 * (size_t)(lo + k)];
                stage[kStageRecs + k] = rec[2 * (size_t)(lo + k) + 1];
            }
        }
        if (AHEAD && row < 8 && active) {
            next = row_fetch(w, (row + 1) % 3 - 1, (row + 1) / 3 - 1, cut, cp, key_base, start, end);
            next_mask = masks[(size_t)(row + 1) * n + i];
        }
        if (staged) wave_lds_sync();
        if (has) {
            if (re - rb <= kRowMaskBits) {
                unsigned long long mm = mrow;
                // the particle itself is in its own row's mask; it adds nothing
                if (i >= rb && i < re) mm &= ~(1ull << (i - rb));
                if (staged)
                    walk_row_mask<1>(A, mm, rb - lo, stage, stage + kStageRecs, pos, vel, dpi.y, h, s);
                else
                    walk_row_mask<2>(A, mm, rb, rec, rec + 1, pos, vel, dpi.y, h, s);
            } else {
                for (uint32_t j = rb; j < re; j++)
                    pair_term_rp(A, i, j, pos, vel, dpi.y, ld_xyz(sp + j), ld_xyz(sv + j), rho_p[j], h, s, hhu);
            }
        }
        if (staged) wave_lds_sync();  // every lane's reads of this row before the next row's stores
    }
}

// Per-lane row table of the flat force schedule, in LDS: for each of the 9
// stencil rows, the sorted index of its first particle and its neighbour
// mask (0 for a culled row).  The flat schedule takes rows of at most 32
// candidates, so a mask is one 32-bit word: 72 bytes per lane, which keeps
// eight 256-lane workgroups' tables within a CU's LDS.
constexpr uint32_t kFlatRowBits = 32;
struct FlatRows {
    uint32_t* rb;  // [9][lanes]
    uint32_t* mm;  // [9][lanes]
    uint32_t lanes, lane;
    __device__ __forceinline__ uint32_t& b(int r) { return rb[r * lanes + lane]; }
    __device__ __forceinline__ uint32_t& m(int r) { return mm[r * lanes + lane]; }
};

// Pressure + viscosity + walls (+ body contact) of sorted particle i (i < n),
// written to vout[id[i]] (and pout[id[i]] when integrating).
//
// Two schedules give bit-identical sums (rows in order, neighbours of a row
// in ascending sorted order):
//  * row loop: the 9 rows in turn; a wavefront waits at every row for its
//    lane with the most neighbours in that row.  Given an LDS slice
//    (`stage`) and a full wavefront, each row's records are staged there
//    first (forces_rows_staged);
//  * flat (MASKS with a row table): the rows with neighbours are gathered
//    first, compacted in row order, then each lane walks its own neighbour
//    stream across them, two neighbours per iteration with their loads
//    issued together, so a wavefront runs for its lane with the most
//    neighbours overall rather than, row by row, for the lane with the
//    most neighbours in that row.  The next row's entry is read from LDS
//    one row ahead.  A particle with a row longer than kFlatRowBits takes
//    the row loop.
template <bool MASKS, bool COUPLED = false, bool DEFER = false, bool AHEAD = false>
__device__ __forceinline__ void forces_particle(uint32_t i, const float4* __restrict__ sp,
                                                const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                                                const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                                                float4* __restrict__ vout, float4* __restrict__ pout,
                                                const uint32_t* __restrict__ id, uint32_t n, float m, float mu,
                                                const fg_common_fluid_params& cp, const fg_kernels_cache& kc,
                                                uint32_t key_base, const unsigned long long* __restrict__ masks,
                                                const Contacts& bc, FlatRows* flat = nullptr,
                                                const DeferredContact* defer = nullptr,
                                                float4* __restrict__ stage = nullptr) {
    const float4 pos = sp[i];
    const float4 vel = sv[i];
    const bool active = pos.w > 0.0f;
    F3 a = f3(0.0f, 0.0f, 0.0f);
    PairAcc A{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    // The staged row loop needs every lane of the wavefront (the run of
    // each row is a wavefront reduction); otherwise the rows are read from
    // memory below.
    const bool staged_rows = MASKS && stage != nullptr && flat == nullptr && __ballot(1) == ~0ull;
    if (staged_rows)
        forces_rows_staged<AHEAD>(A, i, active, pos, vel, active ? rho_p[i] : make_float2(1.0f, 0.0f), sp, sv, rho_p,
                                  start, end, n, cp, key_base, masks, stage);
    if (active) {
        const float h = cp.kernel_radius, hh = h * h;
        const float s = cp.simulation_scale;
        const float cut = hh * 1.0001f;
        const float2 dpi = rho_p[i];
        const RowWalker w = make_rows(pos, cp);
        const float4* rec = MASKS ? nbr_records(masks, n) : nullptr;
        bool rowloop = !staged_rows;
        if (MASKS && flat) {
            // The rows with neighbours, compacted in row order: (first sorted
            // index, mask without the particle itself).
            bool fits = true;
            uint32_t nrows = 0;
#pragma unroll
            for (int row = 0; row < 9; row++) {
                const int dz = row / 3 - 1, dy = row % 3 - 1;
                uint32_t rb = 0, re = 0;
                if (row_range(w, dy, dz, cut, cp, key_base, start, end, n, rb, re)) {
                    if (re - rb <= kFlatRowBits) {
                        unsigned long long m64 = masks[(size_t)row * n + i];
                        // the particle itself is in its own row's mask; it adds nothing
                        if (i >= rb && i < re) m64 &= ~(1ull << (i - rb));
                        const uint32_t mm = (uint32_t)m64;
                        if (mm) {
                            flat->b(nrows) = rb;
                            flat->m(nrows) = mm;
                            nrows++;
                        }
                    } else {
                        fits = false;
                    }
                }
            }
            if (fits) {
                rowloop = false;
                // cm/cb: the current row's remaining mask and first index;
                // nm/nb: the next row's, read from LDS one row ahead.
                uint32_t k = 0;
                uint32_t cm = nrows > 0 ? flat->m(0) : 0u, cb = nrows > 0 ? flat->b(0) : 0u;
                uint32_t nm = nrows > 1 ? flat->m(1) : 0u, nb = nrows > 1 ? flat->b(1) : 0u;
                // The next neighbour of the stream (cm != 0 on entry).
                auto next = [&]() -> uint32_t {
                    const uint32_t j = cb + (uint32_t)__builtin_ctz(cm);
                    cm &= cm - 1u;
                    if (cm == 0u) {
                        cm = nm;
                        cb = nb;
                        k++;
                        nm = k + 1 < nrows ? flat->m(k + 1) : 0u;
                        nb = k + 1 < nrows ? flat->b(k + 1) : 0u;
                    }
                    return j;
                };
                while (cm != 0u) {
                    const uint32_t j1 = next();
                    const bool two = cm != 0u;
                    const uint32_t j2 = two ? next() : j1;
                    const float4 a1 = rec[2 * (size_t)j1], b1 = rec[2 * (size_t)j1 + 1];
                    const float4 a2 = rec[2 * (size_t)j2], b2 = rec[2 * (size_t)j2 + 1];
                    const float3 q1 = make_float3(a1.x, a1.y, a1.z), q2 = make_float3(a2.x, a2.y, a2.z);
                    const float3 v1 = make_float3(b1.x, b1.y, b1.z), v2 = make_float3(b2.x, b2.y, b2.z);
                    const float2 d1 = make_float2(a1.w, b1.w), d2 = make_float2(a2.w, b2.w);
                    asm volatile("" ::"v"(q2.x), "v"(q2.y), "v"(q2.z), "v"(v2.x), "v"(v2.y), "v"(v2.z), "v"(d2.x),
                                 "v"(d2.y));
                    pair_term_masked(A, pos, vel, dpi.y, q1, v1, d1.x, d1.y, h, s);
                    if (two) pair_term_masked(A, pos, vel, dpi.y, q2, v2, d2.x, d2.y, h, s);
                }
            }
        }
        if (rowloop) forces_rowloop<MASKS>(A, i, pos, vel, dpi, sp, sv, rho_p, start, end, n, cp, key_base, masks);
        // pressure: (s d_u / r) = d_u / r_u, so no scale factor remains
        const float iri = 1.0f / dpi.x;
        const float cP = -kc.kSpikyGradient * iri;
        const float cV = kc.kViscosityLaplacian * mu * iri;
        const F3 internal = f3(__builtin_fmaf(cP, A.px, cV * A.vx), __builtin_fmaf(cP, A.py, cV * A.vy),
                               __builtin_fmaf(cP, A.pz, cV * A.vz));
        if (DEFER && bc.coupled && contact_possible(pos, bc)) {
            defer_particle(*defer, i, internal);
            return;
        }
        a = DEFER ? finish_accel<true, true>(internal, m, pos, vel, cp, bc)
                  : finish_accel<COUPLED>(internal, m, pos, vel, cp, bc);
    }
    write_back(id[i], n, pos, vel, a, cp, vout, pout);
}

}  // namespace fg
