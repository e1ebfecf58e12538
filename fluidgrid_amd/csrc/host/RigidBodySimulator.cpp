// RigidBodySimulator.cpp -- host side of the fluid/rigid-body coupling
// (Simulators/RigidBodySimulator.cpp of the reference, restricted to dynamic
// bodies coupled to one fluid).
#include "fluidgrid/RigidBodySimulator.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "fluidgrid/MeshSampler.h"
#include "rigid_math.hpp"

namespace Simulation {
namespace RigidBody {

RigidBodySimulator* RigidBodySimulator::mInstance = nullptr;

namespace {
void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void fg_check(const char* what) {
    if (fg_peek_error() != FG_OK) {
        std::string msg = std::string(what) + ": " + fg_last_error_string();
        (void)fg_last_error();
        throw std::runtime_error(msg);
    }
}

template <class T>
T* dalloc(size_t count) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, sizeof(T) * (count ? count : 1)), "hipMalloc");
    return static_cast<T*>(p);
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

int key_bits_for(unsigned int nb) {
    int b = 1;
    while (b < 32 && ((nb - 1u) >> b) != 0) b++;
    return b;
}


fg_float3 abi3(float3 v) { return fg_float3{v.x, v.y, v.z}; }
}  // namespace

RigidBody::RigidBody(bool isStatic, float mass, float particleRadius, std::vector<float4> worldSpheres)
    : mIsStatic(isStatic), mMass(mass), mParticleRadius(particleRadius), mSpheres(std::move(worldSpheres)) {}

RigidBody* RigidBody::loadFromFile(const std::string& filename, float3 origin) {
    std::ifstream f(filename);
    if (!f) throw std::runtime_error("RigidBody::loadFromFile: cannot open " + filename);
    auto next = [&](const char* what) {
        std::string line;
        if (!std::getline(f, line))
            throw std::runtime_error(std::string("RigidBody::loadFromFile: no ") + what + " line in " + filename);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        return line;
    };
    const bool isStatic = next("type") == "static";
    const float mass = isStatic ? 0.0f : (float)std::atof(next("mass").c_str());
    const float radius = (float)std::atof(next("radius").c_str());
    const std::string mesh = next("mesh");
    const double count = std::atof(next("count").c_str());
    if (!(count >= 0.0) || count > 2147483647.0)
        throw std::runtime_error("RigidBody::loadFromFile: bad sphere count in " + filename);
    std::vector<float4> s;
    s.reserve((size_t)count);
    for (int i = 0; i < (int)count; i++) {
        std::istringstream line(next("sphere"));
        std::string x, y, z;
        line >> x >> y >> z;
        s.push_back(make_float4((float)std::atof(x.c_str()) + origin.x, (float)std::atof(y.c_str()) + origin.y,
                                (float)std::atof(z.c_str()) + origin.z, radius));
    }
    RigidBody* rb = new RigidBody(isStatic, mass, radius, std::move(s));
    rb->setMeshName(mesh);
    return rb;
}

void RigidBody::saveToFile(const std::string& filename, float3 origin) const {
    FILE* f = std::fopen(filename.c_str(), "w");
    if (!f) throw std::runtime_error("RigidBody::saveToFile: cannot open " + filename);
    if (mIsStatic)
        std::fprintf(f, "static\n");
    else
        std::fprintf(f, "dynamic\n%.9g\n", (double)mMass);
    std::fprintf(f, "%.9g\n%s\n%zu\n", (double)mParticleRadius, mMeshName.c_str(), mSpheres.size());
    for (const float4& p : mSpheres)
        std::fprintf(f, "%.9g %.9g %.9g\n", (double)(p.x - origin.x), (double)(p.y - origin.y),
                     (double)(p.z - origin.z));
    if (std::fclose(f) != 0) throw std::runtime_error("RigidBody::saveToFile: write failed for " + filename);
}

RigidBody* RigidBody::fromMesh(const std::string& objFile, bool isStatic, float mass, float particleRadius,
                               float3 origin, bool shellOnly) {
    std::vector<float4> s = sampleMeshSpheres(loadObj(objFile), 2.0f * particleRadius, shellOnly);
    for (float4& p : s) p = make_float4(p.x + origin.x, p.y + origin.y, p.z + origin.z, particleRadius);
    RigidBody* rb = new RigidBody(isStatic, isStatic ? 0.0f : mass, particleRadius, std::move(s));
    rb->setMeshName(objFile);
    return rb;
}

RigidBodySimulator* RigidBodySimulator::instance() {
    if (!mInstance) mInstance = new RigidBodySimulator();
    return mInstance;
}

void RigidBodySimulator::destroy() {
    delete mInstance;
    mInstance = nullptr;
}

RigidBodySimulator::~RigidBodySimulator() { release(); }

void RigidBodySimulator::addElement(RigidBody* rb) {
    if (rb->isStatic())
        mStatic.push_back(rb);
    else
        mDynamic.push_back(rb);
}

// RigidBodySimulator.cpp:146-214 (updateStaticElements): static spheres keyed
// with the reference hash over st_bucket_size buckets, sorted, and the cell
// ranges and sorted positions kept for the whole simulation.
void RigidBodySimulator::buildStaticGrid() {
    const Application::CommonRigidBodyParams& rp =
        Application::ApplicationData::instance()->getCommonRigidBodyParams();
    const Application::CommonFluidParams& fp = Application::ApplicationData::instance()->getCommonFluidParams();
    std::vector<float4> world;
    mStaticRadius = 0.0f;
    for (RigidBody* b : mStatic) {
        for (const float4& p : b->getWorldSpheres()) world.push_back(make_float4(p.x, p.y, p.z, 0.0f));
        mStaticRadius = std::max(mStaticRadius, b->getParticleRadius());
    }
    mStaticCount = (unsigned)world.size();
    mStBuckets = (unsigned)rp.getStNumBuckets();
    if (mStaticCount == 0) return;
    if (mStBuckets == 0) throw std::invalid_argument("RigidBodySimulator: st_bucket_size is empty");
    const unsigned M = mStaticCount;
    float4* d_src = dalloc<float4>(M);
    float4* d_tmp = dalloc<float4>(M);
    float4* d_zero = dalloc<float4>(1);
    float4* d_ident = dalloc<float4>(1);
    unsigned *d_k = dalloc<unsigned>(M), *d_i = dalloc<unsigned>(M);
    d_stSorted = dalloc<float4>(M);
    d_stStart = dalloc<unsigned>(mStBuckets);
    d_stEnd = dalloc<unsigned>(mStBuckets);
    d_stBounds = dalloc<float4>(2);
    const float4 zero = make_float4(0, 0, 0, 0), ident = make_float4(0, 0, 0, 1);
    hip_check(hipMemcpy(d_src, world.data(), sizeof(float4) * M, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_zero, &zero, sizeof(float4), hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_ident, &ident, sizeof(float4), hipMemcpyHostToDevice), "upload");
    hip_check(hipMemset(d_stEnd, 0, sizeof(unsigned) * mStBuckets), "memset");
    const size_t wsBytes = fg_sort_workspace_size(M, key_bits_for(mStBuckets));
    void* ws = nullptr;
    hip_check(hipMalloc(&ws, wsBytes ? wsBytes : 256), "hipMalloc");
    void* prev = fg_get_stream();
    fg_set_stream(mStream);
    // identity pose: the world positions are kept exactly
    gCalcHashKeyWithOffset(d_k, d_i, reinterpret_cast<fg_float4*>(d_zero), reinterpret_cast<fg_float4*>(d_ident),
                           reinterpret_cast<fg_float4*>(d_src), reinterpret_cast<fg_float4*>(d_tmp), M,
                           abi3(rp.getCellSize()), (int)mStBuckets);
    fg_sort_pairs_ws(d_k, d_i, M, key_bits_for(mStBuckets), ws, wsBytes, 1);
    gFindCellOffsetOrderingPos(d_stStart, d_stEnd, d_k, d_i, reinterpret_cast<fg_float4*>(d_tmp),
                               reinterpret_cast<fg_float4*>(d_stSorted), M, mStBuckets);
    // the static box serves the fluid contact (fluid radius 2*pointRadius,
    // margin 0.003) and the body contact (body sphere radius, margin 0.001)
    const float margin = std::max(2.0f * fp.pointRadius, mSphereRadius) + mStaticRadius + 0.003f;
    fg_sphere_bounds(reinterpret_cast<fg_float4*>(d_src), M, margin * 1.01f, reinterpret_cast<fg_float4*>(d_stBounds));
    fg_set_stream(prev);
    fg_check("static body grid");
    hip_check(hipStreamSynchronize(mStream), "sync");
    (void)hipFree(ws);
    dfree(d_src); dfree(d_tmp); dfree(d_zero); dfree(d_ident); dfree(d_k); dfree(d_i);
}

fg_body_grid RigidBodySimulator::staticGrid(float fluidRadius) const {
    const Application::CommonRigidBodyParams& rp =
        Application::ApplicationData::instance()->getCommonRigidBodyParams();
    const Application::CommonFluidParams& fp = Application::ApplicationData::instance()->getCommonFluidParams();
    const Application::SimulationParams& sp = Application::ApplicationData::instance()->getSimulationParams();
    fg_body_grid g{};
    g.orderedSpheres = reinterpret_cast<const fg_float4*>(d_stSorted);
    g.bodyVelocities = nullptr;
    g.gridStart = d_stStart;
    g.gridEnd = d_stEnd;
    g.bounds = reinterpret_cast<const fg_float4*>(d_stBounds);
    g.cellSize = abi3(rp.getCellSize());
    g.numBuckets = mStBuckets;
    g.relation = fg_int3{fp.fluid_to_srb_relation.x, fp.fluid_to_srb_relation.y, fp.fluid_to_srb_relation.z};
    g.numSpheres = mStaticCount;
    g.numBodies = 0;
    g.fluidRadius = fluidRadius;
    g.sphereRadius = mStaticRadius;
    g.kSpring = -sp.springForce;
    g.kDamping = -sp.dampingForce;
    return g;
}

void RigidBodySimulator::release() {
    dfree(d_rel); dfree(d_world); dfree(d_sorted); dfree(d_cm); dfree(d_vel); dfree(d_orient); dfree(d_L);
    dfree(d_force); dfree(d_torque); dfree(d_inv); dfree(d_inertia); dfree(d_pf); dfree(d_pt); dfree(d_bounds);
    dfree(d_mass); dfree(d_sumF); dfree(d_sumT); dfree(d_massd); dfree(d_cmd); dfree(d_inertiad);
    dfree(d_offsets); dfree(d_key); dfree(d_id); dfree(d_start); dfree(d_end);
    dfree(d_stSorted); dfree(d_stBounds); dfree(d_stStart); dfree(d_stEnd); dfree(d_bodyBounds);
    mStaticCount = 0;
    if (d_sortWs) (void)hipFree(d_sortWs);
    d_sortWs = nullptr;
}

void RigidBodySimulator::reset() {
    release();
    const Application::CommonRigidBodyParams& rp =
        Application::ApplicationData::instance()->getCommonRigidBodyParams();
    const size_t B = mDynamic.size();
    std::vector<float4> world;
    mOffsets.assign(1, 0);
    mMass.clear();
    mSphereRadius = 0.0f;
    for (size_t b = 0; b < B; b++) {
        for (const float4& p : mDynamic[b]->getWorldSpheres()) world.push_back(make_float4(p.x, p.y, p.z, (float)b));
        mOffsets.push_back((unsigned)world.size());
        mMass.push_back(mDynamic[b]->getMass());
        mSphereRadius = std::max(mSphereRadius, mDynamic[b]->getParticleRadius());
    }
    const size_t M = world.size();
    mBuckets = (unsigned)rp.getDnNumBuckets();
    buildStaticGrid();
    if (B == 0) return;
    if (mBuckets == 0) throw std::invalid_argument("RigidBodySimulator: dn_bucket_size is empty");

    d_world = dalloc<float4>(M);
    d_sorted = dalloc<float4>(M);
    d_rel = dalloc<float4>(M);
    d_pf = dalloc<float4>(M);
    d_pt = dalloc<float4>(M);
    d_key = dalloc<unsigned>(M);
    d_id = dalloc<unsigned>(M);
    d_start = dalloc<unsigned>(mBuckets);
    d_end = dalloc<unsigned>(mBuckets);
    d_offsets = dalloc<unsigned>(B + 1);
    d_cm = dalloc<float4>(B);
    d_vel = dalloc<float4>(B);
    d_orient = dalloc<float4>(B);
    d_L = dalloc<float4>(B);
    d_force = dalloc<float4>(B);
    d_torque = dalloc<float4>(B);
    d_inv = dalloc<float4>(3 * B);
    d_inertia = dalloc<float4>(3 * B);
    d_mass = dalloc<float>(B);
    d_massd = dalloc<double>(B);
    d_cmd = dalloc<double>(3 * B);
    d_inertiad = dalloc<double>(9 * B);
    d_sumF = dalloc<double>(4 * B);
    d_sumT = dalloc<double>(4 * B);
    d_bounds = dalloc<float4>(2);
    if (B >= 2) d_bodyBounds = dalloc<float4>(2 * B);
    mSortWsBytes = fg_sort_workspace_size((unsigned)M, key_bits_for(mBuckets));
    hip_check(hipMalloc(&d_sortWs, mSortWsBytes ? mSortWsBytes : 256), "hipMalloc");

    hip_check(hipMemcpy(d_world, world.data(), sizeof(float4) * M, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_offsets, mOffsets.data(), sizeof(unsigned) * (B + 1), hipMemcpyHostToDevice), "upload");
    std::vector<double> md(mMass.begin(), mMass.end());
    hip_check(hipMemcpy(d_massd, md.data(), sizeof(double) * B, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_mass, mMass.data(), sizeof(float) * B, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemset(d_end, 0, sizeof(unsigned) * mBuckets), "memset");

    // Centre of mass and inertia tensor in FP64 on the GPU.
    void* prev = fg_get_stream();
    fg_set_stream(mStream);
    fg_body_mass_properties_f64(reinterpret_cast<fg_float4*>(d_world), d_offsets, (unsigned)B, d_massd, d_cmd,
                                d_inertiad);
    fg_set_stream(prev);
    fg_check("fg_body_mass_properties_f64");
    hip_check(hipStreamSynchronize(mStream), "sync");
    mCM.resize(3 * B);
    mInertia.resize(9 * B);
    hip_check(hipMemcpy(mCM.data(), d_cmd, sizeof(double) * 3 * B, hipMemcpyDeviceToHost), "download");
    hip_check(hipMemcpy(mInertia.data(), d_inertiad, sizeof(double) * 9 * B, hipMemcpyDeviceToHost), "download");

    std::vector<float4> cm(B), inv(3 * B), inert(3 * B), zero(B, make_float4(0, 0, 0, 0)),
        orient(B, make_float4(0, 0, 0, 1));
    for (size_t b = 0; b < B; b++) {
        cm[b] = make_float4((float)mCM[3 * b], (float)mCM[3 * b + 1], (float)mCM[3 * b + 2], 0.0f);
        double iv[9];
        fg_host::invert3(&mInertia[9 * b], iv);
        for (int r = 0; r < 3; r++) {
            inv[3 * b + r] = make_float4((float)iv[3 * r], (float)iv[3 * r + 1], (float)iv[3 * r + 2], 0.0f);
            const double* I = &mInertia[9 * b];
            inert[3 * b + r] = make_float4((float)I[3 * r], (float)I[3 * r + 1], (float)I[3 * r + 2], 0.0f);
        }
    }
    // Sphere positions relative to the centre of mass (body frame at rest).
    mRel.resize(M);
    for (size_t b = 0; b < B; b++)
        for (unsigned i = mOffsets[b]; i < mOffsets[b + 1]; i++)
            mRel[i] = make_float4((float)((double)world[i].x - mCM[3 * b]), (float)((double)world[i].y - mCM[3 * b + 1]),
                                  (float)((double)world[i].z - mCM[3 * b + 2]), (float)b);
    hip_check(hipMemcpy(d_rel, mRel.data(), sizeof(float4) * M, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_cm, cm.data(), sizeof(float4) * B, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_inv, inv.data(), sizeof(float4) * 3 * B, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_inertia, inert.data(), sizeof(float4) * 3 * B, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_orient, orient.data(), sizeof(float4) * B, hipMemcpyHostToDevice), "upload");
    for (float4* p : {d_vel, d_L, d_force, d_torque})
        hip_check(hipMemcpy(p, zero.data(), sizeof(float4) * B, hipMemcpyHostToDevice), "upload");
}

void RigidBodySimulator::startInteraction() {
    if (mInteraction) throw std::logic_error("RigidBodySimulator: startInteraction called twice");
    mInteraction = true;
}

void RigidBodySimulator::endInteraction() {
    if (!mInteraction) throw std::logic_error("RigidBodySimulator: endInteraction without startInteraction");
    mInteraction = false;
}

// RigidBodySimulator.cpp:451-474: world transform + key, sort, cell ranges.
void RigidBodySimulator::generateCollisionData() {
    const unsigned M = (unsigned)mRel.size(), B = (unsigned)mMass.size();
    if (M == 0) return;
    const Application::ApplicationData* ad = Application::ApplicationData::instance();
    const Application::CommonRigidBodyParams& rp = const_cast<Application::ApplicationData*>(ad)->getCommonRigidBodyParams();
    const Application::CommonFluidParams& fp = const_cast<Application::ApplicationData*>(ad)->getCommonFluidParams();
    gCalcHashKeyWithOffset(d_key, d_id, reinterpret_cast<fg_float4*>(d_cm), reinterpret_cast<fg_float4*>(d_orient),
                           reinterpret_cast<fg_float4*>(d_rel), reinterpret_cast<fg_float4*>(d_world), M,
                           abi3(rp.getCellSize()), (int)mBuckets);
    fg_sort_pairs_ws(d_key, d_id, M, key_bits_for(mBuckets), d_sortWs, mSortWsBytes, 1);
    gFindCellOffsetOrderingPos(d_start, d_end, d_key, d_id, reinterpret_cast<fg_float4*>(d_world),
                               reinterpret_cast<fg_float4*>(d_sorted), M, mBuckets);
    const float margin = 2.0f * fp.pointRadius + mSphereRadius + 0.003f;
    fg_sphere_bounds(reinterpret_cast<fg_float4*>(d_world), M, margin * 1.01f, reinterpret_cast<fg_float4*>(d_bounds));
    // Broad phase of the body-body contact (RigidBodySimulator.cpp:334-406 on
    // the host in the reference): per-body boxes grown by the contact distance.
    if (B >= 2)
        fg_body_bounds(reinterpret_cast<fg_float4*>(d_world), d_offsets, B, (2.0f * mSphereRadius + 0.001f) * 1.01f,
                       reinterpret_cast<fg_float4*>(d_bodyBounds));
    fg_check("generateCollisionData");
}

// RigidBodySimulator.cpp:503-520 (the reference's body is commented out):
// per sphere, penalty force from the fluid grid and torque about the centre
// of mass; spheres in body order so the sums of applyForces are contiguous.
void RigidBodySimulator::calculateRigidBodyForcesFromFluid(float4* pos, float4* vel, unsigned int* start,
                                                           unsigned int* end, unsigned int numFluid) {
    const unsigned M = (unsigned)mRel.size();
    if (M == 0) return;
    Application::ApplicationData* ad = Application::ApplicationData::instance();
    const Application::CommonRigidBodyParams& rp = ad->getCommonRigidBodyParams();
    const Application::CommonFluidParams& fp = ad->getCommonFluidParams();
    const Application::SimulationParams& sp = ad->getSimulationParams();
    fg_body_forces_from_fluid(reinterpret_cast<fg_float4*>(d_world), reinterpret_cast<fg_float4*>(d_vel),
                              reinterpret_cast<fg_float4*>(d_cm), reinterpret_cast<fg_float4*>(d_pf),
                              reinterpret_cast<fg_float4*>(d_pt), M, (unsigned)mMass.size(), mSphereRadius,
                              reinterpret_cast<fg_float4*>(pos), reinterpret_cast<fg_float4*>(vel), start, end,
                              numFluid, 2.0f * fp.pointRadius, abi3(fp.cell_size), (unsigned)fp.bucket_size.x *
                              fp.bucket_size.y * fp.bucket_size.z,
                              fg_int3{rp.rb_to_fluid_relation.x, rp.rb_to_fluid_relation.y,
                                      rp.rb_to_fluid_relation.z},
                              -sp.springForce, -sp.dampingForce);
    calculateRigidBodyForcesFromRB();
    fg_check("calculateRigidBodyForcesFromFluid");
}

// RigidBodySimulator.cpp:481-501: body-body and body-static contact of every
// dynamic sphere (stencil (1,1,1), margin 0.001), added to the fluid's part
// of the per-sphere forces and torques.
void RigidBodySimulator::calculateRigidBodyForcesFromRB() {
    const unsigned M = (unsigned)mRel.size(), B = (unsigned)mMass.size();
    if (M == 0 || (B < 2 && mStaticCount == 0)) return;
    fg_body_grid dyn = bodyGrid(0.0f), st = staticGrid(0.0f);
    dyn.relation = st.relation = fg_int3{1, 1, 1};
    fg_body_forces_from_bodies(reinterpret_cast<fg_float4*>(d_world), reinterpret_cast<fg_float4*>(d_vel),
                               reinterpret_cast<fg_float4*>(d_cm), reinterpret_cast<fg_float4*>(d_pf),
                               reinterpret_cast<fg_float4*>(d_pt), M, B, mSphereRadius, B >= 2 ? &dyn : nullptr,
                               B >= 2 ? reinterpret_cast<fg_float4*>(d_bodyBounds) : nullptr,
                               mStaticCount ? &st : nullptr, 1);
}

// RigidBodySimulator.cpp:522-566: per-body sums (FP64 here).
void RigidBodySimulator::applyForces() {
    const unsigned B = (unsigned)mMass.size();
    if (B == 0) return;
    fg_body_reduce_f64(reinterpret_cast<fg_float4*>(d_pf), reinterpret_cast<fg_float4*>(d_pt), d_offsets, B,
                       reinterpret_cast<fg_double4*>(d_sumF), reinterpret_cast<fg_double4*>(d_sumT),
                       reinterpret_cast<fg_float4*>(d_force), reinterpret_cast<fg_float4*>(d_torque));
    fg_check("applyForces");
}

// RigidBodySimulator.cpp:568-611 (without the host readback).
void RigidBodySimulator::integrate() {
    const unsigned B = (unsigned)mMass.size();
    if (B == 0) return;
    const Application::CommonRigidBodyParams& rp =
        Application::ApplicationData::instance()->getCommonRigidBodyParams();
    gIntegrateRB(reinterpret_cast<fg_float4*>(d_cm), d_mass, reinterpret_cast<fg_float4*>(d_vel),
                 reinterpret_cast<fg_float4*>(d_force), B, rp.deltaTimeIntegration,
                 fg_float4{rp.global_external_force.x, rp.global_external_force.y, rp.global_external_force.z,
                           rp.global_external_force.w},
                 abi3(rp.volume_min), abi3(rp.volume_max), reinterpret_cast<fg_float4*>(d_L),
                 reinterpret_cast<fg_float4*>(d_torque), reinterpret_cast<fg_float4*>(d_inertia),
                 reinterpret_cast<fg_float4*>(d_orient), reinterpret_cast<fg_float4*>(d_inv), rp.velocity_limit);
    // RigidBodySimulator.cpp:591-592: accumulators are cleared after use.
    fg_memset_u32(reinterpret_cast<unsigned*>(d_force), 0u, 4 * (size_t)B);
    fg_memset_u32(reinterpret_cast<unsigned*>(d_torque), 0u, 4 * (size_t)B);
    fg_check("integrate");
}

fg_body_grid RigidBodySimulator::bodyGrid(float fluidRadius) const {
    const Application::CommonRigidBodyParams& rp =
        Application::ApplicationData::instance()->getCommonRigidBodyParams();
    const Application::CommonFluidParams& fp = Application::ApplicationData::instance()->getCommonFluidParams();
    const Application::SimulationParams& sp = Application::ApplicationData::instance()->getSimulationParams();
    fg_body_grid g{};
    g.orderedSpheres = reinterpret_cast<const fg_float4*>(d_sorted);
    g.bodyVelocities = reinterpret_cast<const fg_float4*>(d_vel);
    g.gridStart = d_start;
    g.gridEnd = d_end;
    g.bounds = reinterpret_cast<const fg_float4*>(d_bounds);
    g.cellSize = abi3(rp.getCellSize());
    g.numBuckets = mBuckets;
    g.relation = fg_int3{fp.fluid_to_srb_relation.x, fp.fluid_to_srb_relation.y, fp.fluid_to_srb_relation.z};
    g.numSpheres = (unsigned)mRel.size();
    g.numBodies = (unsigned)mMass.size();
    g.fluidRadius = fluidRadius;
    g.sphereRadius = mSphereRadius;
    g.kSpring = -sp.springForce;
    g.kDamping = -sp.dampingForce;
    return g;
}

void RigidBodySimulator::downloadBodies(float4* pos, float4* vel, float4* orient, float4* L) const {
    const size_t B = mMass.size();
    if (!B) return;
    hip_check(hipDeviceSynchronize(), "sync");
    if (pos) hip_check(hipMemcpy(pos, d_cm, sizeof(float4) * B, hipMemcpyDeviceToHost), "download");
    if (vel) hip_check(hipMemcpy(vel, d_vel, sizeof(float4) * B, hipMemcpyDeviceToHost), "download");
    if (orient) hip_check(hipMemcpy(orient, d_orient, sizeof(float4) * B, hipMemcpyDeviceToHost), "download");
    if (L) hip_check(hipMemcpy(L, d_L, sizeof(float4) * B, hipMemcpyDeviceToHost), "download");
}

void RigidBodySimulator::uploadBodies(const float4* pos, const float4* vel, const float4* orient, const float4* L) {
    const size_t B = mMass.size();
    if (!B) return;
    hip_check(hipDeviceSynchronize(), "sync");
    if (pos) hip_check(hipMemcpy(d_cm, pos, sizeof(float4) * B, hipMemcpyHostToDevice), "upload");
    if (vel) hip_check(hipMemcpy(d_vel, vel, sizeof(float4) * B, hipMemcpyHostToDevice), "upload");
    if (orient) hip_check(hipMemcpy(d_orient, orient, sizeof(float4) * B, hipMemcpyHostToDevice), "upload");
    if (L) hip_check(hipMemcpy(d_L, L, sizeof(float4) * B, hipMemcpyHostToDevice), "upload");
}

void RigidBodySimulator::downloadSums(double* f4, double* t4) const {
    const size_t B = mMass.size();
    if (!B) return;
    hip_check(hipDeviceSynchronize(), "sync");
    if (f4) hip_check(hipMemcpy(f4, d_sumF, sizeof(double) * 4 * B, hipMemcpyDeviceToHost), "download");
    if (t4) hip_check(hipMemcpy(t4, d_sumT, sizeof(double) * 4 * B, hipMemcpyDeviceToHost), "download");
}

void RigidBodySimulator::downloadMassProperties(double* cm3, double* inertia9) const {
    for (size_t k = 0; k < mCM.size(); k++) cm3[k] = mCM[k];
    for (size_t k = 0; k < mInertia.size(); k++) inertia9[k] = mInertia[k];
}

}  // namespace RigidBody
}  // namespace Simulation
