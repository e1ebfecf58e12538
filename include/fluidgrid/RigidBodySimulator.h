// RigidBodySimulator.h -- rigid bodies coupled to the fluid, after the
// reference's Simulators/RigidBodySimulator.h:130-165 and RigidBody.h:
// dynamic bodies move under the fluid's forces; static bodies are a fixed
// sphere grid the fluid collides with.
//
// The reference wires this path only in commented code
// (HeterogeneousSimulator.cpp:57, 82-102, 113-126); here it is live:
//   generateCollisionData()              body spheres to world space, keyed
//                                        and sorted into the body grid
//   calculateRigidBodyForcesFromFluid()  penalty force and torque per sphere
//                                        from the fluid, then from the other
//                                        bodies and the static spheres
//   applyForces()                        per-body sums in FP64
//   integrate()                          IntegrateRBK per body
// Centre of mass and inertia tensor are computed in FP64 on the GPU from the
// sphere set (equal point masses); the reference hard-codes
// diag(200 m / 12) and its inverse 0.00125 (RigidBody.cpp:116-122).
#pragma once

#include <string>
#include <vector>

#include "ApplicationData.h"

namespace Simulation {
namespace RigidBody {

// RigidBody.h:28-33
struct sInertiaTensor {
    float4 Ix, Iy, Iz;
};

// One body as a set of spheres (RigidBody.h:45-88 without the scene node).
class RigidBody {
public:
    // Spheres in world space; w is ignored.
    RigidBody(bool isStatic, float mass, float particleRadius, std::vector<float4> worldSpheres);

    // RigidBody.cpp:55-108 text format, one item per line: "static", or any
    // other first line followed by the mass; the sphere radius; the name of
    // the source mesh; the sphere count; one "x y z" line per sphere
    // (positions relative to the body origin, placed at `origin`).  Numbers
    // are read as atof does (a malformed number reads as 0); a file that
    // ends early throws std::runtime_error.
    static RigidBody* loadFromFile(const std::string& filename, float3 origin);
    // The same format, written so that loadFromFile reads back the same
    // floats (9 significant digits); positions relative to `origin`.
    void saveToFile(const std::string& filename, float3 origin) const;
    // A body sampled from a Wavefront OBJ mesh (sampleMeshSpheres with
    // spacing 2 * particleRadius, solid or shell), translated by `origin`.
    static RigidBody* fromMesh(const std::string& objFile, bool isStatic, float mass, float particleRadius,
                               float3 origin, bool shellOnly);

    const std::vector<float4>& getWorldSpheres() const { return mSpheres; }
    int getNumParticles() const { return (int)mSpheres.size(); }
    float getParticleRadius() const { return mParticleRadius; }
    float getMass() const { return mMass; }
    bool isStatic() const { return mIsStatic; }
    const std::string& getMeshName() const { return mMeshName; }
    void setMeshName(const std::string& name) { mMeshName = name; }

private:
    bool mIsStatic;
    float mMass;
    float mParticleRadius;
    std::vector<float4> mSpheres;
    std::string mMeshName;
};

class RigidBodySimulator {
public:
    static RigidBodySimulator* instance();
    static void destroy();

    void addElement(RigidBody* rbElement);
    // Uploads the bodies, computes centre of mass and inertia (FP64) and
    // allocates the per-step buffers (RigidBodySimulator.cpp:216-311).
    void reset();
    void release();

    void startInteraction();
    void endInteraction();
    void generateCollisionData();
    void calculateRigidBodyForcesFromFluid(float4* dfluidParticlePos, float4* dfluidParticleVel,
                                           unsigned int* dfluidHashStart, unsigned int* dfluidHashEnd,
                                           unsigned int numFluid);
    // Body-body and body-static contact added to the per-sphere forces (run
    // by calculateRigidBodyForcesFromFluid; a no-op with one body and no
    // static spheres).
    void calculateRigidBodyForcesFromRB();
    void applyForces();
    void integrate();

    int getNumDynamicBodies() const { return (int)mMass.size(); }
    int getNumDynamicSpheres() const { return (int)mRel.size(); }
    int getNumStaticSpheres() const { return (int)mStaticCount; }
    float4* getDynamicVelocities() { return d_vel; }
    // The dynamic-body grid for the coupled fluid force pass.
    fg_body_grid bodyGrid(float fluidRadius) const;
    // The static-body grid (RigidBodySimulator.cpp:146-214): built once at
    // reset(), REF_XOR keys over st_bucket_size buckets, no velocities.
    fg_body_grid staticGrid(float fluidRadius) const;

    // Extensions: state readback (positions, velocities, orientations,
    // angular momenta: float4 per body) and the FP64 sums of the last step.
    void downloadBodies(float4* pos, float4* vel, float4* orient, float4* L) const;
    // The inverse: restores the per-body state (each array B float4s).
    void uploadBodies(const float4* pos, const float4* vel, const float4* orient, const float4* L);
    void downloadSums(double* forces4, double* torques4) const;
    void downloadMassProperties(double* cm3, double* inertia9) const;
    void setStream(hipStream_t s) { mStream = s; }

private:
    RigidBodySimulator() = default;
    ~RigidBodySimulator();

    std::vector<RigidBody*> mDynamic;
    std::vector<RigidBody*> mStatic;
    std::vector<float4> mRel;          // spheres relative to the centre of mass, w = body
    std::vector<unsigned int> mOffsets;
    std::vector<float> mMass;
    std::vector<double> mCM, mInertia;  // FP64, 3 and 9 per body
    float mSphereRadius = 0.0f;
    bool mInteraction = false;
    hipStream_t mStream = nullptr;

    // device
    float4 *d_rel = nullptr, *d_world = nullptr, *d_sorted = nullptr, *d_cm = nullptr, *d_vel = nullptr;
    float4 *d_orient = nullptr, *d_L = nullptr, *d_force = nullptr, *d_torque = nullptr, *d_inv = nullptr;
    float4 *d_inertia = nullptr, *d_pf = nullptr, *d_pt = nullptr, *d_bounds = nullptr;
    float* d_mass = nullptr;
    double *d_sumF = nullptr, *d_sumT = nullptr, *d_massd = nullptr, *d_cmd = nullptr, *d_inertiad = nullptr;
    unsigned int *d_offsets = nullptr, *d_key = nullptr, *d_id = nullptr, *d_start = nullptr, *d_end = nullptr;
    void* d_sortWs = nullptr;
    size_t mSortWsBytes = 0;
    unsigned int mBuckets = 0;
    float4* d_bodyBounds = nullptr;  // [2B] broad phase of the body-body contact
    // static bodies
    float4 *d_stSorted = nullptr, *d_stBounds = nullptr;
    unsigned int *d_stStart = nullptr, *d_stEnd = nullptr;
    unsigned int mStaticCount = 0, mStBuckets = 0;
    float mStaticRadius = 0.0f;
    void buildStaticGrid();

    static RigidBodySimulator* mInstance;
};

}  // namespace RigidBody
}  // namespace Simulation
