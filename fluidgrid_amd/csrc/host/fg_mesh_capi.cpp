// fg_mesh_capi.cpp -- C-ABI over the body file format and the mesh sampler.
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "fluidgrid/MeshSampler.h"
#include "fluidgrid/RigidBodySimulator.h"
#include "fluidgrid_amd_mesh.h"

using namespace Simulation::RigidBody;

namespace {
int fail(int code, const std::string& msg) {
    fg_record_error(code, msg.c_str());
    return code;
}

TriangleMesh mesh_of(const fg_float3* v, unsigned nv, const unsigned* t, unsigned nt) {
    TriangleMesh m;
    m.vertices.resize(nv);
    for (unsigned i = 0; i < nv; i++) m.vertices[i] = make_float3(v[i].x, v[i].y, v[i].z);
    m.triangles.assign(t, t + 3 * (size_t)nt);
    for (uint32_t i : m.triangles)
        if (i >= nv) throw std::invalid_argument("triangle corner beyond the vertex count");
    return m;
}

bool bad_mesh(const fg_float3* v, unsigned nv, const unsigned* t, unsigned nt) {
    return (nv && !v) || (nt && !t) || nv == 0;
}
}  // namespace

extern "C" {

int fg_mesh_load_obj(const char* path, unsigned int* numVertices, unsigned int* numTriangles, fg_float3* vertices,
                     unsigned int* triangles) {
    if (!path || !numVertices || !numTriangles)
        return fail(FG_ERR_INVALID_ARGUMENT, "fg_mesh_load_obj: null pointer");
    try {
        const TriangleMesh m = loadObj(path);
        const size_t nt = m.numTriangles();
        if (m.vertices.size() > 0xFFFFFFFFu || nt > 0xFFFFFFFFu)
            return fail(FG_ERR_INVALID_ARGUMENT, "fg_mesh_load_obj: more than 2^32 vertices or triangles");
        if ((vertices || triangles) && (*numVertices < m.vertices.size() || *numTriangles < nt)) {
            *numVertices = (unsigned)m.vertices.size();
            *numTriangles = (unsigned)nt;
            return fail(FG_ERR_INVALID_ARGUMENT, "fg_mesh_load_obj: arrays smaller than the mesh");
        }
        if (vertices)
            for (size_t i = 0; i < m.vertices.size(); i++)
                vertices[i] = fg_float3{m.vertices[i].x, m.vertices[i].y, m.vertices[i].z};
        if (triangles) std::memcpy(triangles, m.triangles.data(), sizeof(unsigned) * m.triangles.size());
        *numVertices = (unsigned)m.vertices.size();
        *numTriangles = (unsigned)nt;
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(FG_ERR_INVALID_ARGUMENT, e.what());
    }
}

int fg_mesh_lattice(const fg_float3* vertices, unsigned int numVertices, float spacing, fg_float3* origin,
                    unsigned int* dims) {
    if (!vertices || !numVertices || !origin || !dims)
        return fail(FG_ERR_INVALID_ARGUMENT, "fg_mesh_lattice: null pointer or no vertices");
    try {
        const MeshLattice L = meshLattice(mesh_of(vertices, numVertices, nullptr, 0), spacing);
        *origin = fg_float3{L.origin.x, L.origin.y, L.origin.z};
        dims[0] = L.nx;
        dims[1] = L.ny;
        dims[2] = L.nz;
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(FG_ERR_INVALID_ARGUMENT, e.what());
    }
}

int fg_mesh_winding_numbers(const fg_float3* vertices, unsigned int numVertices, const unsigned int* triangles,
                            unsigned int numTriangles, float spacing, float* winding) {
    if (!winding || bad_mesh(vertices, numVertices, triangles, numTriangles))
        return fail(FG_ERR_INVALID_ARGUMENT, "fg_mesh_winding_numbers: null pointer or no vertices");
    TriangleMesh m;
    MeshLattice L;
    try {
        m = mesh_of(vertices, numVertices, triangles, numTriangles);
        L = meshLattice(m, spacing);
    } catch (const std::exception& e) {
        return fail(FG_ERR_INVALID_ARGUMENT, std::string("fg_mesh_winding_numbers: ") + e.what());
    }
    try {
        const std::vector<float> w = windingNumbers(m, L);
        std::memcpy(winding, w.data(), sizeof(float) * w.size());
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(FG_ERR_HIP, e.what());
    }
}

int fg_mesh_sample_spheres(const fg_float3* vertices, unsigned int numVertices, const unsigned int* triangles,
                           unsigned int numTriangles, float spacing, int shellOnly, fg_float4* out,
                           unsigned int capacity, unsigned int* count) {
    if (!count || bad_mesh(vertices, numVertices, triangles, numTriangles))
        return fail(FG_ERR_INVALID_ARGUMENT, "fg_mesh_sample_spheres: null pointer or no vertices");
    TriangleMesh m;
    try {
        m = mesh_of(vertices, numVertices, triangles, numTriangles);
        (void)meshLattice(m, spacing);
    } catch (const std::exception& e) {
        return fail(FG_ERR_INVALID_ARGUMENT, std::string("fg_mesh_sample_spheres: ") + e.what());
    }
    try {
        const std::vector<float4> s = sampleMeshSpheres(m, spacing, shellOnly != 0);
        if (s.size() > 0xFFFFFFFFu) return fail(FG_ERR_INVALID_ARGUMENT, "fg_mesh_sample_spheres: over 2^32 points");
        *count = (unsigned)s.size();
        if (!out) return FG_OK;
        if (capacity < s.size()) return fail(FG_ERR_INVALID_ARGUMENT, "fg_mesh_sample_spheres: capacity too small");
        for (size_t i = 0; i < s.size(); i++) out[i] = fg_float4{s[i].x, s[i].y, s[i].z, s[i].w};
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(FG_ERR_HIP, e.what());
    }
}

int fg_body_file_write(const char* path, int isStatic, float mass, float particleRadius, const char* meshName,
                       const fg_float4* relativeSpheres, unsigned int numSpheres) {
    if (!path || (numSpheres && !relativeSpheres))
        return fail(FG_ERR_INVALID_ARGUMENT, "fg_body_file_write: null pointer");
    const std::string mesh = meshName ? meshName : "";
    if (mesh.find_first_of("\r\n") != std::string::npos)
        return fail(FG_ERR_INVALID_ARGUMENT, "fg_body_file_write: mesh name spans lines");
    try {
        std::vector<float4> s(numSpheres);
        for (unsigned i = 0; i < numSpheres; i++)
            s[i] = make_float4(relativeSpheres[i].x, relativeSpheres[i].y, relativeSpheres[i].z, particleRadius);
        RigidBody rb(isStatic != 0, isStatic ? 0.0f : mass, particleRadius, std::move(s));
        rb.setMeshName(mesh);
        rb.saveToFile(path, make_float3(0.0f, 0.0f, 0.0f));
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(FG_ERR_INVALID_ARGUMENT, e.what());
    }
}

int fg_body_file_read(const char* path, int* isStatic, float* mass, float* particleRadius, unsigned int* numSpheres,
                      fg_float4* spheres, unsigned int capacity, char* meshName, size_t meshNameLen) {
    if (!path || !isStatic || !mass || !particleRadius || !numSpheres)
        return fail(FG_ERR_INVALID_ARGUMENT, "fg_body_file_read: null pointer");
    try {
        std::unique_ptr<RigidBody> rb(RigidBody::loadFromFile(path, make_float3(0.0f, 0.0f, 0.0f)));
        const std::vector<float4>& s = rb->getWorldSpheres();
        if (spheres && capacity < s.size())
            return fail(FG_ERR_INVALID_ARGUMENT, "fg_body_file_read: capacity too small");
        *isStatic = rb->isStatic() ? 1 : 0;
        *mass = rb->getMass();
        *particleRadius = rb->getParticleRadius();
        *numSpheres = (unsigned)s.size();
        if (spheres)
            for (size_t i = 0; i < s.size(); i++) spheres[i] = fg_float4{s[i].x, s[i].y, s[i].z, s[i].w};
        if (meshName && meshNameLen) {
            const std::string& n = rb->getMeshName();
            const size_t k = n.size() < meshNameLen - 1 ? n.size() : meshNameLen - 1;
            std::memcpy(meshName, n.data(), k);
            meshName[k] = '\0';
        }
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(FG_ERR_INVALID_ARGUMENT, e.what());
    }
}

}  // extern "C"
