// MeshSampler.cpp -- OBJ reading, the sampling lattice and the inside test
// of triangle meshes (the GPU part is fg_mesh_winding_lattice).
#include "fluidgrid/MeshSampler.h"

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "fluidgrid_amd.h"

namespace Simulation {
namespace RigidBody {

namespace {
void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct DeviceBuffer {
    void* p = nullptr;
    explicit DeviceBuffer(size_t bytes) { hip_check(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc"); }
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

// One OBJ face corner ("i", "i/t", "i//n", "i/t/n") to a 0-based index.
uint32_t obj_corner(const std::string& tok, size_t numVertices, const std::string& filename, int line) {
    const std::string idx = tok.substr(0, tok.find('/'));
    char* end = nullptr;
    const long v = std::strtol(idx.c_str(), &end, 10);
    const long i = v > 0 ? v - 1 : (long)numVertices + v;  // negative: relative to the last vertex read
    if (idx.empty() || *end != '\0' || v == 0 || i < 0 || i >= (long)numVertices)
        throw std::runtime_error("loadObj: bad face corner '" + tok + "' at " + filename + ":" + std::to_string(line));
    return (uint32_t)i;
}

#pragma clang fp contract(off)
// The kernel's lattice point: product and sum each rounded to FP32.
float3 lattice_point(const MeshLattice& L, uint32_t ix, uint32_t iy, uint32_t iz) {
    return make_float3(L.origin.x + (float)ix * L.spacing, L.origin.y + (float)iy * L.spacing,
                       L.origin.z + (float)iz * L.spacing);
}
#pragma clang fp contract(fast)
}  // namespace

TriangleMesh loadObj(const std::string& filename) {
    std::ifstream f(filename);
    if (!f) throw std::runtime_error("loadObj: cannot open " + filename);
    TriangleMesh m;
    std::string text;
    int line = 0;
    while (std::getline(f, text)) {
        line++;
        std::istringstream in(text);
        std::string op;
        if (!(in >> op)) continue;
        if (op == "v") {
            double x, y, z;
            if (!(in >> x >> y >> z))
                throw std::runtime_error("loadObj: bad vertex at " + filename + ":" + std::to_string(line));
            m.vertices.push_back(make_float3((float)x, (float)y, (float)z));
        } else if (op == "f") {
            std::vector<uint32_t> c;
            std::string tok;
            while (in >> tok) c.push_back(obj_corner(tok, m.vertices.size(), filename, line));
            if (c.size() < 3)
                throw std::runtime_error("loadObj: face with fewer than 3 corners at " + filename + ":" +
                                         std::to_string(line));
            for (size_t k = 1; k + 1 < c.size(); k++) {
                m.triangles.push_back(c[0]);
                m.triangles.push_back(c[k]);
                m.triangles.push_back(c[k + 1]);
            }
        }
    }
    return m;
}

MeshLattice meshLattice(const TriangleMesh& mesh, float spacing) {
    if (!(spacing > 0.0f)) throw std::runtime_error("meshLattice: spacing must be positive");
    if (mesh.vertices.empty()) throw std::runtime_error("meshLattice: mesh has no vertices");
    float lo[3], hi[3];
    for (int a = 0; a < 3; a++) lo[a] = hi[a] = (&mesh.vertices[0].x)[a];
    for (const float3& v : mesh.vertices)
        for (int a = 0; a < 3; a++) {
            const float c = (&v.x)[a];
            lo[a] = std::fmin(lo[a], c);
            hi[a] = std::fmax(hi[a], c);
        }
    uint32_t n[3];
    float o[3];
    for (int a = 0; a < 3; a++) {
        const double cnt = std::ceil(((double)hi[a] - (double)lo[a]) / (double)spacing);
        if (cnt > 65536.0) throw std::runtime_error("meshLattice: more than 65536 points along an axis");
        n[a] = cnt < 1.0 ? 1u : (uint32_t)cnt;
        o[a] = (float)(0.5 * ((double)lo[a] + (double)hi[a]) - 0.5 * (double)(n[a] - 1) * (double)spacing);
    }
    return MeshLattice{make_float3(o[0], o[1], o[2]), spacing, n[0], n[1], n[2]};
}

std::vector<float> windingNumbers(const TriangleMesh& mesh, const MeshLattice& L) {
    const size_t nt = mesh.numTriangles();
    std::vector<float> w(L.size());
    if (w.empty()) return w;
    DeviceBuffer dv(sizeof(float3) * mesh.vertices.size()), dt(sizeof(uint32_t) * 3 * nt),
        dw(sizeof(float) * w.size());
    hip_check(hipMemcpy(dv.p, mesh.vertices.data(), sizeof(float3) * mesh.vertices.size(), hipMemcpyHostToDevice),
              "windingNumbers");
    hip_check(hipMemcpy(dt.p, mesh.triangles.data(), sizeof(uint32_t) * 3 * nt, hipMemcpyHostToDevice),
              "windingNumbers");
    const int rc = fg_mesh_winding_lattice(static_cast<const fg_float3*>(dv.p), (unsigned)mesh.vertices.size(),
                                           static_cast<const unsigned*>(dt.p), (unsigned)nt,
                                           fg_float3{L.origin.x, L.origin.y, L.origin.z}, L.spacing, L.nx, L.ny,
                                           L.nz, static_cast<float*>(dw.p));
    if (rc != FG_OK) {
        std::string msg = std::string("windingNumbers: ") + fg_last_error_string();
        (void)fg_last_error();
        throw std::runtime_error(msg);
    }
    hip_check(fg_synchronize() == FG_OK ? hipSuccess : hipErrorUnknown, "windingNumbers: synchronize");
    hip_check(hipMemcpy(w.data(), dw.p, sizeof(float) * w.size(), hipMemcpyDeviceToHost), "windingNumbers");
    return w;
}

std::vector<float4> sampleMeshSpheres(const TriangleMesh& mesh, float spacing, bool shellOnly) {
    const MeshLattice L = meshLattice(mesh, spacing);
    const std::vector<float> w = windingNumbers(mesh, L);
    auto inside = [&](int64_t x, int64_t y, int64_t z) {
        if (x < 0 || y < 0 || z < 0 || x >= L.nx || y >= L.ny || z >= L.nz) return false;
        return w[(size_t)x + (size_t)L.nx * ((size_t)y + (size_t)L.ny * (size_t)z)] > 0.5f;
    };
    std::vector<float4> out;
    for (int64_t z = 0; z < L.nz; z++)
        for (int64_t y = 0; y < L.ny; y++)
            for (int64_t x = 0; x < L.nx; x++) {
                if (!inside(x, y, z)) continue;
                if (shellOnly && inside(x - 1, y, z) && inside(x + 1, y, z) && inside(x, y - 1, z) &&
                    inside(x, y + 1, z) && inside(x, y, z - 1) && inside(x, y, z + 1))
                    continue;
                const float3 p = lattice_point(L, (uint32_t)x, (uint32_t)y, (uint32_t)z);
                out.push_back(make_float4(p.x, p.y, p.z, 0.5f * spacing));
            }
    return out;
}

}  // namespace RigidBody
}  // namespace Simulation
