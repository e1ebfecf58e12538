// fg_sph_tile.hpp -- LDS-tiled density and force passes (LINEAR keys, fast
// math).
//
// A tile is the kTileLanes (128) consecutive sorted particles of one
// workgroup.
// Their cells lie in one or two x-rows of cells (a row is the run of cells at
// fixed (cy, cz); LINEAR keys make its cells consecutive keys, so its
// particles are consecutive in sorted order).  The 27-cell stencils of a
// row's cells [cxa, cxb] are covered by 9 contiguous sorted ranges, one per
// neighbour row (cy+dy, cz+dz), cells [cxa-1, cxb+1]: the tile's segments.
// Both passes stage the segments in LDS with direct global-to-LDS loads; a
// particle's range in a stencil row (the cells of row_range) is a sub-range
// of its segment.
//
//  * density: the row loop of density_particle, reading candidates from LDS
//    (same groups of four, same order: the same bits).  Besides (rho, p) and
//    the neighbour record it writes the particle's neighbours as a list of
//    LDS indices (u16, at most kListMax, row order, ascending within a row,
//    the particle itself left out) instead of the 9 row masks;
//  * forces: the same tile layout, with the neighbour records staged; each
//    lane walks its own list, two neighbours per iteration, so a wavefront
//    runs for its lane with the most neighbours instead of waiting, row by
//    row, for the lane with the most neighbours in that row.  Pairs are
//    evaluated in the mask order of forces_particle: the same bits.
//
// A tile whose particles span more than two rows, or a segment more than
// kTileSpanMax cells, or whose segments hold more than kTileCap particles,
// takes the per-lane global-memory path (density_particle / forces_particle
// with row masks) for every particle; so does a particle with more than
// kListMax neighbours or a stencil row longer than 64 candidates (its list
// count is kListOverflow and its row masks are written).  Both passes make
// the same decision from the same grid, so they agree.
#pragma once
#include "fg_sph_fast.hpp"

namespace fg {

constexpr int kTileLanes = 128;
constexpr int kTileRowsMax = 2;
constexpr int kTileSegs = 9 * kTileRowsMax;
constexpr int kTileSpanMax = 128;  // cells per segment: two 64-lane chunks
constexpr uint32_t kTileCap = 1664;  // 2 x 1664 x 16 B of records: three force workgroups per CU
constexpr uint32_t kListMax = 64;
constexpr uint16_t kListOverflow = 0xFFFFu;

// Neighbour buffer (fg_neighbor_mask_bytes): 9 row masks per particle, then
// from 256-byte boundaries the 32-byte neighbour records, a u16 list count
// per particle and the lists (kListMax u16 per particle, in 16-byte groups).
__host__ __device__ __forceinline__ size_t nbr_align256(size_t x) { return (x + 255) & ~(size_t)255; }
__host__ __device__ __forceinline__ size_t nbr_count_offset(uint32_t n) {
    return nbr_align256(nbr_record_offset(n) + 2 * 16 * (size_t)n);
}
__host__ __device__ __forceinline__ size_t nbr_list_offset(uint32_t n) {
    return nbr_align256(nbr_count_offset(n) + 2 * (size_t)n);
}
// After the lists, one 256-byte descriptor per tile (TileDesc), written by
// the tiled density pass for the force pass.
constexpr uint32_t kTileDescWords = 64;
__host__ __device__ __forceinline__ size_t nbr_desc_offset(uint32_t n) {
    return nbr_align256(nbr_list_offset(n) + 2 * (size_t)kListMax * n);
}
// After the descriptors, the FP16 cell-offset pair records of the mixed
// precision form (fg_sph_mixed.hpp): 16 bytes per particle and two padding
// records.
__host__ __device__ __forceinline__ size_t nbr_half_offset(uint32_t n) {
    const size_t tiles = ((size_t)n + kTileLanes - 1) / kTileLanes;
    return nbr_align256(nbr_desc_offset(n) + tiles * kTileDescWords * sizeof(uint32_t));
}
__host__ __device__ __forceinline__ size_t nbr_buffer_bytes(uint32_t n) {
    return nbr_half_offset(n) + 16 * ((size_t)n + 2);
}
__device__ __forceinline__ uint16_t* nbr_counts(unsigned long long* masks, uint32_t n) {
    return reinterpret_cast<uint16_t*>(reinterpret_cast<char*>(masks) + nbr_count_offset(n));
}
__device__ __forceinline__ const uint16_t* nbr_counts(const unsigned long long* masks, uint32_t n) {
    return reinterpret_cast<const uint16_t*>(reinterpret_cast<const char*>(masks) + nbr_count_offset(n));
}
// The lists in 16-byte groups of 8 entries: group G of particle i is uint4
// [G * n + i], entries 2p and 2p+1 of the group's pair p its word p (low half
// first).  The force pass fetches a whole list with at most 8 16-byte loads
// issued together; the groups of consecutive particles are contiguous.
__device__ __forceinline__ uint4* nbr_lists(unsigned long long* masks, uint32_t n) {
    return reinterpret_cast<uint4*>(reinterpret_cast<char*>(masks) + nbr_list_offset(n));
}
__device__ __forceinline__ const uint4* nbr_lists(const unsigned long long* masks, uint32_t n) {
    return reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(masks) + nbr_list_offset(n));
}

// A tile's staging plan as the force pass needs it: whether the tile took
// the global path, and each segment's first sorted index, length and LDS
// offset (64 words).
struct TileDesc {
    uint32_t fallback, nrows, total, pad;
    uint32_t sb[kTileSegs], len[kTileSegs], off[kTileSegs];
    uint32_t unused[kTileDescWords - 4 - 3 * kTileSegs];
};
static_assert(sizeof(TileDesc) == kTileDescWords * sizeof(uint32_t), "tile descriptor size");
__device__ __forceinline__ uint32_t* nbr_descs(unsigned long long* masks, uint32_t n) {
    return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(masks) + nbr_desc_offset(n));
}
__device__ __forceinline__ const uint32_t* nbr_descs(const unsigned long long* masks, uint32_t n) {
    return reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(masks) + nbr_desc_offset(n));
}

// Geometry of one tile, in LDS.
struct TileGeom {
    uint32_t fallback;            // (force pass) the tile took the per-lane global path
    uint32_t nrows;               // distinct cell rows among the tile's particles (1 or 2)
    uint32_t split;               // first lane of the second row
    uint32_t total;               // particles staged
    int cy[kTileRowsMax], cz[kTileRowsMax];
    int cell0[kTileRowsMax];      // first cell (x index) of the row's segments
    int span[kTileRowsMax];       // cells per segment
    uint32_t sb[kTileSegs];       // first sorted index of each segment
    uint32_t len[kTileSegs];      // particles in each segment (0: empty or outside the grid)
    uint32_t off[kTileSegs];      // LDS index of each segment's first particle
};

// Per-tile LDS scratch of the set-up: the lanes' rows and cells.
struct TileSetupLds {
    uint32_t row[kTileLanes];
    int cx[kTileLanes];
};

__device__ __forceinline__ uint32_t tile_row_of(const CellXYZ& c, const fg_common_fluid_params& cp) {
    return (uint32_t)c.y + (uint32_t)cp.bucket_size.y * (uint32_t)c.z;
}

// Segment s of the tile: row q = s / 9 of the tile, neighbour row r = s % 9
// with dz = r / 3 - 1, dy = r % 3 - 1 (the row order of the passes).
__device__ __forceinline__ bool tile_segment_row(const TileGeom& g, int s, const fg_common_fluid_params& cp, int& cy,
                                                 int& cz) {
    const int q = s / 9, r = s % 9;
    cy = g.cy[q] + (r % 3 - 1);
    cz = g.cz[q] + (r / 3 - 1);
    return q < (int)g.nrows && cy >= 0 && cz >= 0 && cy < cp.bucket_size.y && cz < cp.bucket_size.z;
}

// Builds the tile's geometry and segments; every lane of the
// workgroup calls it (it synchronises).  Lane t holds sorted particle
// i = tile * kTileLanes + t when t < nvalid (have).  Returns nonzero when the
// tile takes the per-lane global path (the same value in every lane).
//
// Latency matters here more than work: each wavefront issues the cell-start
// and cell-end loads of all its segments at once (one memory round trip) and
// finds each segment's bounds from two ballots,
// every lane derives the rows' geometry from broadcast LDS reads instead of
// waiting for one lane to, and one wavefront scans the segment offsets.
constexpr int kSegsPerWave = (kTileSegs + kTileLanes / 64 - 1) / (kTileLanes / 64);

__device__ __forceinline__ uint32_t tile_setup(TileGeom& g, TileSetupLds& L, uint32_t n, uint32_t nvalid, bool have,
                                               const float4& p, const fg_common_fluid_params& cp, uint32_t key_base,
                                               const uint32_t* __restrict__ start,
                                               const uint32_t* __restrict__ end) {
    const int t = threadIdx.x;
    if (have) {
        const CellXYZ c = linear_cell(p.x, p.y, p.z, cp);
        L.row[t] = tile_row_of(c, cp);
        L.cx[t] = c.x;
    } else {
        L.row[t] = kEmpty;
        L.cx[t] = 0;
    }
    __syncthreads();
    // Sorted keys make each row's lanes contiguous, and cx non-decreasing
    // within a row.
    const bool trans = have && t > 0 && L.row[t] != L.row[t - 1];
    if (trans) g.split = (uint32_t)t;
    const int ntrans = __syncthreads_count(trans);
    if (ntrans >= kTileRowsMax) return 1u;
    const uint32_t last = nvalid - 1;
    const uint32_t split = ntrans == 0 ? last + 1 : g.split;
    const uint32_t nrows = (uint32_t)ntrans + 1;
    int cy[kTileRowsMax], cz[kTileRowsMax], cell0[kTileRowsMax], span[kTileRowsMax];
    bool wide = false;
#pragma unroll
    for (int q = 0; q < kTileRowsMax; q++) {
        const uint32_t a = q == 0 ? 0u : split, b = (uint32_t)q + 1 < nrows ? split - 1 : last;
        const uint32_t row = L.row[q < (int)nrows ? a : 0u];
        cy[q] = (int)(row % (uint32_t)cp.bucket_size.y);
        cz[q] = (int)(row / (uint32_t)cp.bucket_size.y);
        const int lo = max(L.cx[q < (int)nrows ? a : 0u] - 1, 0);
        const int hi = min(L.cx[q < (int)nrows ? b : 0u] + 1, cp.bucket_size.x - 1);
        cell0[q] = lo;
        span[q] = hi - lo + 1;
        wide = wide || (q < (int)nrows && span[q] > kTileSpanMax);
    }
    if (wide) return 1u;
    if (t == 0) {
        g.nrows = nrows;
        g.split = split;
#pragma unroll
        for (int q = 0; q < kTileRowsMax; q++) {
            g.cy[q] = cy[q];
            g.cz[q] = cz[q];
            g.cell0[q] = cell0[q];
            g.span[q] = span[q];
        }
    }
    // One wavefront per segment (segments w, w+4, ...), one lane per cell of
    // each 64-cell chunk; all of a wavefront's loads are issued before any
    // is used.
    const int wave = t >> 6, lane = t & 63;
    const int nseg = 9 * (int)nrows;
    uint32_t v[kSegsPerWave][2], e[kSegsPerWave][2];
    static_assert(kTileRowsMax == 2, "the selects below pick one of two tile rows");
    static_assert(kTileSpanMax == 128, "two 64-cell chunks per segment");
#pragma unroll
    for (int j = 0; j < kSegsPerWave; j++) {
        const int s = wave + j * (kTileLanes / 64);
        const bool q1 = s >= 9;
        const int r = q1 ? s - 9 : s;
        const int ry = (q1 ? cy[1] : cy[0]) + (r % 3 - 1), rz = (q1 ? cz[1] : cz[0]) + (r / 3 - 1);
        const bool inside = s < nseg && ry >= 0 && rz >= 0 && ry < cp.bucket_size.y && rz < cp.bucket_size.z;
        const int sp_ = q1 ? span[1] : span[0];
        const uint32_t key0 = key_base + linear_key(q1 ? cell0[1] : cell0[0], ry, rz, cp);
#pragma unroll
        for (int c = 0; c < 2; c++) {
            v[j][c] = kEmpty;
            e[j][c] = 0u;
            if (inside && lane + 64 * c < sp_) {
                v[j][c] = start[key0 + (uint32_t)(lane + 64 * c)];
                e[j][c] = end[key0 + (uint32_t)(lane + 64 * c)];  // stale for an empty cell; only occupied ends are used
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kSegsPerWave; j++) {
        const int s = wave + j * (kTileLanes / 64);
        if (s >= nseg) break;
        const unsigned long long occ0 = __ballot(v[j][0] != kEmpty), occ1 = __ballot(v[j][1] != kEmpty);
        // the segment's first particle: the start of its first occupied
        // cell; one past its last: the end of its last occupied cell
        uint32_t sb = 0u, se = 0u;
        if (occ0 | occ1) {
            se = min(occ1 ? (uint32_t)__shfl((int)e[j][1], 63 - __builtin_clzll(occ1))
                          : (uint32_t)__shfl((int)e[j][0], 63 - __builtin_clzll(occ0)),
                     n);
            sb = occ0 ? (uint32_t)__shfl((int)v[j][0], __builtin_ctzll(occ0))
                      : (uint32_t)__shfl((int)v[j][1], __builtin_ctzll(occ1));
        }
        if (lane == 0) {
            g.sb[s] = sb;
            g.len[s] = se > sb ? se - sb : 0u;
        }
    }
    __syncthreads();
    // LDS offsets: an exclusive scan of the segment lengths in wavefront 0.
    if (wave == 0) {
        const uint32_t len = lane < nseg ? g.len[lane] : 0u;
        uint32_t inc = len;
#pragma unroll
        for (int d = 1; d < 32; d <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)inc, d);
            if (lane >= d) inc += o;
        }
        if (lane < nseg) g.off[lane] = inc - len;
        if (lane == nseg - 1) g.total = inc;
    }
    __syncthreads();
    return g.total > kTileCap ? 1u : 0u;
}

// Copies each segment's records into LDS (dst[off + k] = src[(sb + k) *
// stride]) with direct global-to-LDS loads: one 16-byte element per lane, the
// instruction's LDS base wave-uniform.  Segments are shared round-robin by
// the wavefronts.  The caller waits (s_waitcnt) and synchronises.
__device__ __forceinline__ void tile_fill(const TileGeom& g, const float4* __restrict__ src, uint32_t stride,
                                          float4* dst) {
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int nseg = 9 * (int)g.nrows;
    for (int s = wave; s < nseg; s += kTileLanes / 64) {
        const uint32_t len = g.len[s], sb = g.sb[s], off = g.off[s];
        for (uint32_t b = 0; b < len; b += 64) {
            if (b + (uint32_t)lane < len)
                __builtin_amdgcn_global_load_lds(
                    (const void*)(src + (size_t)(sb + b + (uint32_t)lane) * stride),
                    (__attribute__((address_space(3))) void*)(dst + off + b), 16, 0, 0);
        }
    }
}

__device__ __forceinline__ void tile_fill_wait() {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// Density of the tile's particle in lane t (LDS mode): density_particle's row
// loop over LDS candidates (row ranges from row_range, staged at
// off[s] + (rb - sb[s])), with the neighbours appended to the lane's list in
// LDS (llist, kListMax entries) and copied out in 16-byte groups at the end
// (or, on overflow, the row masks written).  v is the particle's sorted
// velocity (for its neighbour record).
#pragma clang fp contract(off)
__device__ __forceinline__ void tile_density_particle(const TileGeom& g, const float4* __restrict__ lpos,
                                                      uint16_t* __restrict__ llist, uint32_t i, const float4& p,
                                                      const float4& v, float2* __restrict__ rho_p,
                                                      const uint32_t* __restrict__ start,
                                                      const uint32_t* __restrict__ end, uint32_t n, float k,
                                                      float rho0, float m, const fg_common_fluid_params& cp,
                                                      const fg_kernels_cache& kc,
                                                      unsigned long long* __restrict__ masks) {
    float4* rec = nbr_records(masks, n);
    uint16_t* cnt_out = nbr_counts(masks, n);
    if (!(p.w > 0.0f)) {
        store_record(i, rho_p[i], p, v, rec);
        cnt_out[i] = 0;
        return;
    }
    const int q = (int)threadIdx.x < (int)g.split ? 0 : 1;
    const float s = cp.simulation_scale, s2 = s * s;
    const float h = cp.kernel_radius, hh = h * h;
    const float hu = h / s, hhu = hu * hu, h2u = cp.kernel_radius_squared / s2;
    const float cut = hh * 1.0001f;
    const RowWalker w = make_rows(p, cp);
    const float3 far = make_float3(3.0e38f, 3.0e38f, 3.0e38f);
    float acc = 0.0f;
    uint32_t cnt = 0;
    bool overflow = false;
    unsigned long long msk[9];
#pragma unroll
    for (int row = 0; row < 9; row++) {
        msk[row] = 0ull;
        const int dz = row / 3 - 1, dy = row % 3 - 1;
        uint32_t rb, re;
        if (!row_range(w, dy, dz, cut, cp, 0u, start, end, n, rb, re)) continue;
        const int sg = q * 9 + row;
        const uint32_t len = re - rb, lb = rb - g.sb[sg] + g.off[sg];
        const float4* pp = lpos + lb;
        const uint32_t n4 = len & ~3u;
        unsigned long long mm = 0ull;
        uint32_t kk = 0;
        for (; kk < n4; kk += 4) {
            const float4 a0 = pp[kk], a1 = pp[kk + 1], a2 = pp[kk + 2], a3 = pp[kk + 3];
            const uint32_t nib = density4(p, make_float3(a0.x, a0.y, a0.z), make_float3(a1.x, a1.y, a1.z),
                                          make_float3(a2.x, a2.y, a2.z), make_float3(a3.x, a3.y, a3.z), h2u, hhu,
                                          acc);
            mm |= (unsigned long long)nib << (kk & 63u);
        }
        if (kk < len) {
            const uint32_t rest = len - kk;
            const float4 a0 = pp[kk];
            const float3 t1 = rest > 1 ? make_float3(pp[kk + 1].x, pp[kk + 1].y, pp[kk + 1].z) : far;
            const float3 t2 = rest > 2 ? make_float3(pp[kk + 2].x, pp[kk + 2].y, pp[kk + 2].z) : far;
            const uint32_t nib = density4(p, make_float3(a0.x, a0.y, a0.z), t1, t2, far, h2u, hhu, acc);
            mm |= (unsigned long long)nib << (kk & 63u);
        }
        msk[row] = mm;
        if (len > kRowMaskBits) {
            overflow = true;  // the force pass re-tests this row
            continue;
        }
        // the particle itself adds no force
        if (i >= rb && i < re) mm &= ~(1ull << (i - rb));
        // appended as two 32-bit words, low word first (ascending order)
        uint32_t word = (uint32_t)mm, base = lb;
        for (int half = 0; half < 2; half++) {
            while (word) {
                const uint32_t b = (uint32_t)__builtin_ctz(word);
                word &= word - 1u;
                if (cnt < kListMax) llist[cnt] = (uint16_t)(base + b);
                cnt++;
            }
            word = (uint32_t)(mm >> 32);
            base = lb + 32u;
            if (!word) break;
        }
    }
    if (cnt > kListMax) overflow = true;
    store_density(i, m * kc.kPoly6 * (s2 * s2 * s2) * acc, k, rho0, rho_p, rec, p, v);
    if (overflow) {
#pragma unroll
        for (int row = 0; row < 9; row++) masks[(size_t)row * n + i] = msk[row];
        cnt_out[i] = kListOverflow;
        return;
    }
    cnt_out[i] = (uint16_t)cnt;
    uint4* lists = nbr_lists(masks, n);
    const uint4* l4 = reinterpret_cast<const uint4*>(llist);
#pragma unroll
    for (int G = 0; G < (int)(kListMax / 8); G++)
        if ((uint32_t)(8 * G) < cnt) lists[(size_t)G * n + i] = l4[G];
}
#pragma clang fp contract(fast)

// Pressure + viscosity + walls (+ contact) of the tile's particle i from its
// neighbour list over the LDS records (recA: x, y, z, 1/rho; recB: vx, vy,
// vz, p/rho), or, for an overflowed list, from the row masks over the global
// records (forces_particle's row loop).  l0 and l1 are the list's first two
// groups (entries past cnt are ignored).  Returns the acceleration.
template <bool COUPLED>
__device__ __forceinline__ F3 tile_forces_particle(const float4* __restrict__ recA, const float4* __restrict__ recB,
                                                   uint32_t i, const float4& pos, const float4& vel, float2 dpi,
                                                   uint32_t cnt, uint4 l0, uint4 l1, const float4* __restrict__ sp,
                                                   const float4* __restrict__ sv, const float2* __restrict__ rho_p,
                                                   const uint32_t* __restrict__ start,
                                                   const uint32_t* __restrict__ end, uint32_t n, float m, float mu,
                                                   const fg_common_fluid_params& cp, const fg_kernels_cache& kc,
                                                   const unsigned long long* __restrict__ masks, const Contacts& bc) {
    const float h = cp.kernel_radius;
    const float s = cp.simulation_scale;
    PairAcc A{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (cnt != kListOverflow) {
        // The list in 8-entry groups, each loaded two groups ahead of its
        // use.  A group's 16 record reads are issued together, then its 8
        // pairs evaluated in order (entries past the count read record 0 and
        // are skipped), so the LDS latency is paid once per group.
        const uint4* lists = nbr_lists(masks, n);
        auto load_group = [&](uint32_t G) -> uint4 {
            uint4 q = make_uint4(0u, 0u, 0u, 0u);
            if (G < kListMax / 8u && 8u * G < cnt) q = lists[(size_t)G * n + i];
            return q;
        };
        // (the first two groups were loaded by the caller, before the staging)
        uint4 g0 = l0, g1 = l1;
        for (uint32_t G = 0; 8u * G < cnt; G++) {
            const uint4 cur = g0;
            g0 = g1;
            g1 = load_group(G + 2u);
            const uint32_t k = 8u * G;
            const uint32_t w[4] = {cur.x, cur.y, cur.z, cur.w};
            float4 ra[8], rb[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                uint32_t j = (w[u >> 1] >> (16 * (u & 1))) & 0xFFFFu;
                j = k + (uint32_t)u < cnt ? j : 0u;
                ra[u] = recA[j];
                rb[u] = recB[j];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                if (k + (uint32_t)u < cnt)
                    pair_term_masked(A, pos, vel, dpi.y, make_float3(ra[u].x, ra[u].y, ra[u].z),
                                     make_float3(rb[u].x, rb[u].y, rb[u].z), ra[u].w, rb[u].w, h, s);
            }
        }
    } else {
        forces_rowloop<true>(A, i, pos, vel, dpi, sp, sv, rho_p, start, end, n, cp, 0u, masks);
    }
    const float iri = 1.0f / dpi.x;
    const float cP = -kc.kSpikyGradient * iri;
    const float cV = kc.kViscosityLaplacian * mu * iri;
    const F3 internal = f3(__builtin_fmaf(cP, A.px, cV * A.vx), __builtin_fmaf(cP, A.py, cV * A.vy),
                           __builtin_fmaf(cP, A.pz, cV * A.vz));
    return finish_accel<COUPLED>(internal, m, pos, vel, cp, bc);
}

}  // namespace fg
