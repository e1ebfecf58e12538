/*
 * fluidgrid_amd_mesh.h -- C-ABI over the body file format and the mesh
 * sampler (include/fluidgrid/MeshSampler.h, RigidBody::loadFromFile /
 * saveToFile).  The reference reads body particle files
 * (RigidBody::loadFromFile, Simulators/RigidBody.cpp:55-108) whose sphere
 * sets were sampled from meshes outside the reference; these functions make
 * and read such files.  All arrays are host memory; the winding numbers are
 * evaluated on the GPU (fg_mesh_winding_lattice).  Errors return an FG_ERR_*
 * code with the message in fg_last_error_string().
 */
#ifndef FLUIDGRID_AMD_MESH_H
#define FLUIDGRID_AMD_MESH_H

#include "fluidgrid_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reads a Wavefront OBJ file (MeshSampler.h: loadObj).  Call once with
 * vertices and triangles NULL for the counts, then with arrays of
 * *numVertices and 3 * *numTriangles entries. */
int fg_mesh_load_obj(const char* path, unsigned int* numVertices, unsigned int* numTriangles, fg_float3* vertices,
                     unsigned int* triangles);

/* The sampling lattice of a vertex set (MeshSampler.h: meshLattice):
 * origin and point counts dims[3]. */
int fg_mesh_lattice(const fg_float3* vertices, unsigned int numVertices, float spacing, fg_float3* origin,
                    unsigned int* dims);

/* Winding number of every lattice point (x fastest) into winding
 * (dims[0] * dims[1] * dims[2] floats). */
int fg_mesh_winding_numbers(const fg_float3* vertices, unsigned int numVertices, const unsigned int* triangles,
                            unsigned int numTriangles, float spacing, float* winding);

/* Lattice points inside the mesh (MeshSampler.h: sampleMeshSpheres), w =
 * spacing / 2.  *count receives the number of points; out may be NULL to
 * query it.  With out given and capacity below the count, nothing is
 * written and FG_ERR_INVALID_ARGUMENT is returned. */
int fg_mesh_sample_spheres(const fg_float3* vertices, unsigned int numVertices, const unsigned int* triangles,
                           unsigned int numTriangles, float spacing, int shellOnly, fg_float4* out,
                           unsigned int capacity, unsigned int* count);

/* Writes a body particle file (RigidBody::saveToFile with origin 0):
 * positions relative to the body origin, w ignored; mass is ignored for a
 * static body. */
int fg_body_file_write(const char* path, int isStatic, float mass, float particleRadius, const char* meshName,
                       const fg_float4* relativeSpheres, unsigned int numSpheres);

/* Reads a body particle file (RigidBody::loadFromFile with origin 0):
 * spheres relative to the body origin with w = the radius.  Call with
 * spheres NULL for the count; meshName (meshNameLen bytes, may be NULL)
 * receives the mesh line, truncated and NUL-terminated. */
int fg_body_file_read(const char* path, int* isStatic, float* mass, float* particleRadius, unsigned int* numSpheres,
                      fg_float4* spheres, unsigned int capacity, char* meshName, size_t meshNameLen);

#ifdef __cplusplus
}
#endif

#endif /* FLUIDGRID_AMD_MESH_H */
