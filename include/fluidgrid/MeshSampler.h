// MeshSampler.h -- triangle meshes sampled into rigid-body sphere sets.
//
// The reference's body particle files (RigidBody.cpp:55-108) name the mesh
// they were made from; the sampling itself happened outside the reference.
// Here a Wavefront OBJ mesh is read, a cubic lattice of the requested
// spacing is laid over its bounds, and the lattice points inside the mesh
// (generalized winding number above 1/2, evaluated on the GPU by
// fg_mesh_winding_lattice) become the body's sphere centres, either all of
// them (a solid body) or only those with a face neighbour outside (a shell).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace Simulation {
namespace RigidBody {

struct TriangleMesh {
    std::vector<float3> vertices;
    std::vector<uint32_t> triangles;  // three vertex indices per triangle, counter-clockwise seen from outside
    size_t numTriangles() const { return triangles.size() / 3; }
};

// Wavefront OBJ: "v x y z [w]" vertices and "f" faces whose corners are
// "i", "i/t", "i//n" or "i/t/n", with 1-based or negative (relative)
// indices; polygons are split into a fan around their first corner.  Other
// statements are ignored.  Throws std::runtime_error on a missing file, a
// malformed vertex or a face corner out of range.
TriangleMesh loadObj(const std::string& filename);

// The sampling lattice over a mesh's vertex bounds [lo, hi]: per axis
// n = max(1, ceil((hi - lo) / spacing)) points, centred on the bounds:
// origin = (lo + hi) / 2 - (n - 1) spacing / 2 (FP64, then rounded to FP32).
// Point (ix, iy, iz) is origin + i * spacing per axis in FP32.
struct MeshLattice {
    float3 origin;
    float spacing;
    uint32_t nx, ny, nz;
    uint64_t size() const { return (uint64_t)nx * ny * nz; }
};
MeshLattice meshLattice(const TriangleMesh& mesh, float spacing);

// Winding number of every lattice point (x fastest), computed on the GPU.
std::vector<float> windingNumbers(const TriangleMesh& mesh, const MeshLattice& lattice);

// Lattice points with winding number > 1/2, in lattice order; with
// shellOnly, only those with at least one of their six face neighbours
// outside (or beyond the lattice).  w of each result is spacing / 2.
std::vector<float4> sampleMeshSpheres(const TriangleMesh& mesh, float spacing, bool shellOnly);

}  // namespace RigidBody
}  // namespace Simulation
