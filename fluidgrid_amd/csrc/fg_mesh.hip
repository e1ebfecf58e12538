// fg_mesh.hip -- generalized winding numbers of a triangle mesh on a point
// lattice, the device half of sampling a mesh into the rigid-body sphere sets
// of the reference's body particle files (RigidBody.cpp:55-108 reads a file
// that names its source mesh; the files were made from meshes offline).
//
// w(p) = sum over triangles of Omega_t(p) / (4 pi), Omega_t the signed solid
// angle of triangle t seen from p (Van Oosterom & Strackee, IEEE Trans.
// Biomed. Eng. 30(2), 1983):
//   tan(Omega / 2) = a.(b x c) / (|a||b||c| + (a.b)|c| + (b.c)|a| + (c.a)|b|)
// with a, b, c the corners minus p.  w is 1 inside and 0 outside a closed,
// outward-oriented mesh and degrades gracefully for open or self-intersecting
// ones (Jacobson et al., SIGGRAPH 2013).
//
// Every lattice point is independent and reads every triangle, an all-pairs
// shape: a workgroup stages a tile of triangles in LDS, and each lane then
// evaluates its points against the tile from LDS broadcasts.  The work is
// ALU-bound (a square root per corner and an atan2 per pair); the per-point
// sum is FP64, so adding thousands of terms adds no error of its own.
#include "fg_internal.hpp"

namespace fg {

constexpr int kMeshBlock = 256;
constexpr int kMeshPts = 2;  // lattice points per lane

#pragma clang fp contract(off)
// Lattice point (ix, iy, iz): origin + i * spacing per axis, product and sum
// each rounded to FP32.
__device__ __forceinline__ float3 mesh_lattice_point(fg_float3 o, float s, uint32_t ix, uint32_t iy, uint32_t iz) {
    return make_float3(o.x + (float)ix * s, o.y + (float)iy * s, o.z + (float)iz * s);
}
#pragma clang fp contract(fast)

// atan2 of one triangle's half solid angle seen from p.
__device__ __forceinline__ float half_solid_angle(float3 p, float4 v0, float4 v1, float4 v2) {
    const float ax = v0.x - p.x, ay = v0.y - p.y, az = v0.z - p.z;
    const float bx = v1.x - p.x, by = v1.y - p.y, bz = v1.z - p.z;
    const float cx = v2.x - p.x, cy = v2.y - p.y, cz = v2.z - p.z;
    const float la = sqrtf(ax * ax + ay * ay + az * az);
    const float lb = sqrtf(bx * bx + by * by + bz * bz);
    const float lc = sqrtf(cx * cx + cy * cy + cz * cz);
    const float det = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
    const float den = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + (bx * cx + by * cy + bz * cz) * la +
                      (cx * ax + cy * ay + cz * az) * lb;
    return atan2f(det, den);
}

__global__ void __launch_bounds__(kMeshBlock)
    k_mesh_winding(const fg_float3* __restrict__ verts, uint32_t nverts, const uint32_t* __restrict__ tris,
                   uint32_t ntris, fg_float3 origin, float spacing, uint32_t nx, uint32_t ny, uint64_t npts,
                   float* __restrict__ out) {
    __shared__ float4 tile[3][kMeshBlock];
    const uint64_t base = (uint64_t)blockIdx.x * (kMeshBlock * kMeshPts) + threadIdx.x;
    float3 p[kMeshPts];
    double acc[kMeshPts];
#pragma unroll
    for (int k = 0; k < kMeshPts; k++) {
        const uint64_t i = base + (uint64_t)k * kMeshBlock;
        const uint64_t q = i < npts ? i : 0;
        const uint32_t ix = (uint32_t)(q % nx), iy = (uint32_t)((q / nx) % ny), iz = (uint32_t)(q / ((uint64_t)nx * ny));
        p[k] = mesh_lattice_point(origin, spacing, ix, iy, iz);
        acc[k] = 0.0;
    }
    for (uint32_t t0 = 0; t0 < ntris; t0 += kMeshBlock) {
        const uint32_t t = t0 + threadIdx.x;
        float4 c0 = make_float4(0.f, 0.f, 0.f, 0.f), c1 = c0, c2 = c0;
        if (t < ntris) {
            const uint32_t i0 = tris[3 * (uint64_t)t], i1 = tris[3 * (uint64_t)t + 1], i2 = tris[3 * (uint64_t)t + 2];
            // A triangle naming a vertex that does not exist is left out
            // (w = 0 marks it; it is skipped, not evaluated).
            if (i0 < nverts && i1 < nverts && i2 < nverts) {
                const fg_float3 a = verts[i0], b = verts[i1], c = verts[i2];
                c0 = make_float4(a.x, a.y, a.z, 1.f);
                c1 = make_float4(b.x, b.y, b.z, 0.f);
                c2 = make_float4(c.x, c.y, c.z, 0.f);
            }
        }
        __syncthreads();  // the previous tile has been read by every lane
        tile[0][threadIdx.x] = c0;
        tile[1][threadIdx.x] = c1;
        tile[2][threadIdx.x] = c2;
        __syncthreads();
        const uint32_t cnt = min((uint32_t)kMeshBlock, ntris - t0);
        for (uint32_t j = 0; j < cnt; j++) {
            const float4 v0 = tile[0][j], v1 = tile[1][j], v2 = tile[2][j];
            if (v0.w == 0.f) continue;  // the same j in every lane: a uniform branch
#pragma unroll
            for (int k = 0; k < kMeshPts; k++) acc[k] += (double)half_solid_angle(p[k], v0, v1, v2);
        }
    }
    // w = sum(2 atan2) / (4 pi)
    constexpr double kInv2Pi = 0.15915494309189533577;
#pragma unroll
    for (int k = 0; k < kMeshPts; k++) {
        const uint64_t i = base + (uint64_t)k * kMeshBlock;
        if (i < npts) out[i] = (float)(acc[k] * kInv2Pi);
    }
}

}  // namespace fg

using namespace fg;

extern "C" {

int fg_mesh_winding_lattice(const fg_float3* vertices, unsigned int numVertices, const unsigned int* triangles,
                            unsigned int numTriangles, fg_float3 origin, float spacing, unsigned int nx,
                            unsigned int ny, unsigned int nz, float* winding) {
    const uint64_t npts = (uint64_t)nx * ny * nz;
    if (npts == 0) return FG_OK;
    if (!winding || (numTriangles && (!vertices || !triangles))) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_mesh_winding_lattice: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (!(spacing > 0.0f) || npts > (1ull << 34)) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_mesh_winding_lattice: non-positive spacing or more than 2^34 points");
        return FG_ERR_INVALID_ARGUMENT;
    }
    const State& st = state();
    const uint64_t blocks = (npts + kMeshBlock * kMeshPts - 1) / (kMeshBlock * kMeshPts);
    hipLaunchKernelGGL(k_mesh_winding, dim3((unsigned)blocks), dim3(kMeshBlock), 0, st.stream, vertices, numVertices,
                       triangles, numTriangles, origin, spacing, nx, ny, npts, winding);
    return check_launch("fg_mesh_winding_lattice") ? FG_OK : fg_peek_error();
}

}  // extern "C"
