// BatchedSimulator.cpp -- host side of the batched-scene step (fg_batch.hip).
#include "fluidgrid/BatchedSimulator.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "fg_batch_bodies.hpp"
#include "rigid_math.hpp"

namespace Simulation {
namespace Simulators {

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

int key_bits_for(uint64_t nb) {
    int b = 1;
    while (b < 32 && ((nb - 1u) >> b) != 0) b++;
    return b;
}

// The persistent launches' work queue of a tuning (counters null: plain launches).
fg::WorkQueue work_queue(unsigned* counters, const BatchTuning& t) {
    return fg::WorkQueue{t.queueItems > 0 ? counters : nullptr, (uint32_t)(t.queueItems > 0 ? t.queueItems : 1)};
}

struct StreamScope {
    void* prev;
    explicit StreamScope(hipStream_t s) : prev(fg_get_stream()) { fg_set_stream(s); }
    ~StreamScope() { fg_set_stream(prev); }
};

// Same constants as HeterogeneousSimulator (HeterogeneousSimulator.cpp:33-39).
fg_kernels_cache kernels_for(float h) {
    using Application::PI;
    fg_kernels_cache kc;
    const double hd = h;
    kc.kPoly6 = (float)(315.0f / (64.0f * PI * std::pow(hd, 9.0)));
    kc.kSpikyGradient = (float)(-(45.0f / (PI * std::pow(hd, 6.0))));
    kc.kViscosityLaplacian = (float)(45.0f / (PI * std::pow(hd, 6.0)));
    return kc;
}
}  // namespace

BatchedSimulator::BatchedSimulator(const std::vector<BatchScene>& scenes) {
    if (scenes.empty()) throw std::invalid_argument("BatchedSimulator: no scenes");
    uint64_t n = 0, nb = 0;
    std::vector<unsigned> sceneOf;
    std::vector<float4> pos, vel;
    for (size_t s = 0; s < scenes.size(); s++) {
        const BatchScene& sc = scenes[s];
        const auto& cp = sc.common;
        if (cp.bucket_size.x <= 0 || cp.bucket_size.y <= 0 || cp.bucket_size.z <= 0)
            throw std::invalid_argument("BatchedSimulator: scene " + std::to_string(s) + " has an empty grid");
        if (sc.velocities.size() != sc.positions.size())
            throw std::invalid_argument("BatchedSimulator: positions and velocities differ in length");
        // The density and force passes walk the 27 cells around a particle's
        // own, which hold every neighbour only when each cell edge is at
        // least the kernel radius (in world units, h / simulation_scale).
        const float hw = cp.kernel_radius / cp.simulation_scale;
        auto reach = [hw](float cell) { return std::ceil(hw / cell); };  // the launchers' linear_reach
        if (reach(cp.cell_size.x) > 1.0f || reach(cp.cell_size.y) > 1.0f || reach(cp.cell_size.z) > 1.0f)
            throw UnsupportedScene("BatchedSimulator: scene " + std::to_string(s) +
                                        " has cells smaller than the kernel radius (the 27-cell stencil would miss "
                                        "neighbours)");
        const uint64_t cb = (uint64_t)cp.bucket_size.x * cp.bucket_size.y * cp.bucket_size.z;
        fg_scene_desc d{};
        d.common = Application::abi_cast<fg_common_fluid_params>(cp);
        d.kernels = kernels_for(cp.kernel_radius);
        d.mass = sc.fluid.mass;
        d.gasConstant = sc.fluid.gasConstant;
        d.restDensity = sc.fluid.restDensity;
        d.viscosity = sc.fluid.viscosity;
        d.keyBase = (unsigned)nb;
        d.particleBegin = (unsigned)n;
        d.particleCount = (unsigned)sc.positions.size();
        mDescs.push_back(d);
        n += sc.positions.size();
        nb += cb;
        if (n >= (1u << 30) || nb >= 0xFFFFFFFFull)
            throw std::invalid_argument("BatchedSimulator: too many particles or buckets");
        sceneOf.insert(sceneOf.end(), sc.positions.size(), (unsigned)s);
        pos.insert(pos.end(), sc.positions.begin(), sc.positions.end());
        vel.insert(vel.end(), sc.velocities.begin(), sc.velocities.end());
    }
    mN = (unsigned)n;
    mBuckets = (unsigned)nb;
    mKeyBits = key_bits_for(nb);

    d_descs = dalloc<fg_scene_desc>(mDescs.size());
    d_sceneOf = dalloc<unsigned>(mN);
    d_pos = dalloc<float4>(mN);
    d_vel = dalloc<float4>(mN);
    d_spos = dalloc<float4>(mN);
    d_svel = dalloc<float4>(mN);
    d_rho = dalloc<float2>(mN);
    d_key = dalloc<unsigned>(mN);
    d_id = dalloc<unsigned>(mN);
    d_start = dalloc<unsigned>(mBuckets);
    d_end = dalloc<unsigned>(mBuckets);
    d_masks = static_cast<unsigned long long*>(static_cast<void*>(dalloc<char>(fg_neighbor_mask_bytes(mN))));
    mSortWsBytes = fg_sort_workspace_size(mN, mKeyBits);
    hip_check(hipMalloc(&d_sortWs, mSortWsBytes ? mSortWsBytes : 256), "hipMalloc");

    hip_check(hipMemcpy(d_descs, mDescs.data(), sizeof(fg_scene_desc) * mDescs.size(), hipMemcpyHostToDevice),
              "upload");
    hip_check(hipMemcpy(d_sceneOf, sceneOf.data(), sizeof(unsigned) * mN, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_pos, pos.data(), sizeof(float4) * mN, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_vel, vel.data(), sizeof(float4) * mN, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemset(d_rho, 0, sizeof(float2) * mN), "memset");
    hip_check(hipMemset(d_end, 0, sizeof(unsigned) * mBuckets), "memset");
    d_queue = dalloc<unsigned>(fg::kQueueSlots * fg::kQueueStride);
    hip_check(hipStreamCreateWithFlags(&mStream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamCreateWithFlags(&mBodyStream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : mEv) hip_check(hipEventCreate(&e), "hipEventCreate");
    for (auto& e : mSync) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    int blk = 256, bits = 8, fd = 0, ff = 0;
    fg_get_tuning(&blk, &bits);
    fg_get_schedule(&fd, &ff);
    mTuning.sphBlock = blk;
    mTuning.sortDigitBits = bits;
    mTuning.flatDensity = fd != 0;
    mTuning.flatForces = ff != 0;
    mSceneBodies.resize(scenes.size());
}

BatchedSimulator::~BatchedSimulator() {
    invalidateGraph();
    if (mStream) (void)hipStreamSynchronize(mStream);
    if (mBodyStream) (void)hipStreamSynchronize(mBodyStream);
    for (auto& e : mEv)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : mSync)
        if (e) (void)hipEventDestroy(e);
    if (mBodyStream) (void)hipStreamDestroy(mBodyStream);
    if (mStream) (void)hipStreamDestroy(mStream);
    dfree(d_descs); dfree(d_sceneOf); dfree(d_pos); dfree(d_vel); dfree(d_spos); dfree(d_svel); dfree(d_rho);
    dfree(d_key); dfree(d_id); dfree(d_start); dfree(d_end); dfree(d_masks); dfree(d_queue);
    if (d_sortWs) (void)hipFree(d_sortWs);
    releaseBodies();
}

void BatchedSimulator::releaseBodies() {
    if (d_bscenes) (void)hipFree(d_bscenes);
    d_bscenes = nullptr;
    dfree(d_bodyScene); dfree(d_boffsets); dfree(d_bkey); dfree(d_bid); dfree(d_bstart); dfree(d_bend);
    dfree(d_brel); dfree(d_bworld); dfree(d_bsorted); dfree(d_bpf); dfree(d_bpt); dfree(d_bcm); dfree(d_bvel);
    dfree(d_borient); dfree(d_bL); dfree(d_bforce); dfree(d_btorque); dfree(d_binv); dfree(d_sceneBounds);
    dfree(d_bodyBounds); dfree(d_bmass); dfree(d_bsumF); dfree(d_bsumT);
    dfree(d_defAcc); dfree(d_defList); dfree(d_defCount);
    if (d_bsortWs) (void)hipFree(d_bsortWs);
    d_bsortWs = nullptr;
    mBodiesReady = false;
}

void BatchedSimulator::addBodies(unsigned scene, const BatchBodies& b) {
    if (scene >= mDescs.size()) throw std::invalid_argument("BatchedSimulator::addBodies: no such scene");
    if (mBodiesReady) throw std::logic_error("BatchedSimulator::addBodies: call before the first step");
    if (!mSceneBodies[scene].masses.empty())
        throw std::logic_error("BatchedSimulator::addBodies: scene " + std::to_string(scene) + " already has bodies");
    const size_t B = b.masses.size();
    if (B == 0 || b.offsets.size() != B + 1 || b.offsets[0] != 0 || b.offsets[B] != b.worldSpheres.size())
        throw std::invalid_argument("BatchedSimulator::addBodies: offsets do not match the bodies and spheres");
    for (size_t k = 0; k < B; k++)
        if (b.offsets[k + 1] <= b.offsets[k] || !(b.masses[k] > 0.0))
            throw std::invalid_argument("BatchedSimulator::addBodies: a body without spheres or mass");
    if (b.rigid.getDnNumBuckets() <= 0 || !(b.sphereRadius > 0.0f))
        throw std::invalid_argument("BatchedSimulator::addBodies: empty body grid or sphere radius");
    mSceneBodies[scene] = b;
    invalidateGraph();
}

// The single-fluid RigidBodySimulator::reset, for every scene at once.
void BatchedSimulator::setupBodies() {
    if (mBodiesReady) return;
    const unsigned S = (unsigned)mDescs.size();
    std::vector<float4> world;
    std::vector<unsigned> offsets(1, 0), bodyScene;
    std::vector<float> mass;
    std::vector<fg::BatchBodyScene> bs(S);
    uint64_t nbB = 0;
    for (unsigned s = 0; s < S; s++) {
        const BatchBodies& in = mSceneBodies[s];
        fg::BatchBodyScene& d = bs[s];
        std::memset(static_cast<void*>(&d), 0, sizeof(d));
        d.body_begin = (unsigned)mass.size();
        d.sphere_begin = (unsigned)world.size();
        for (size_t b = 0; b < in.masses.size(); b++) {
            const unsigned gb = (unsigned)mass.size();
            for (unsigned i = in.offsets[b]; i < in.offsets[b + 1]; i++)
                world.push_back(make_float4(in.worldSpheres[i].x, in.worldSpheres[i].y, in.worldSpheres[i].z, (float)gb));
            offsets.push_back((unsigned)world.size());
            mass.push_back((float)in.masses[b]);  // RigidBody keeps a float mass
            bodyScene.push_back(s);
        }
        d.body_end = (unsigned)mass.size();
        d.sphere_end = (unsigned)world.size();
        if (d.body_end == d.body_begin) continue;
        const auto& rp = in.rigid;
        const fg_scene_desc& fd = mDescs[s];
        const auto& cp = fd.common;
        const float r = in.sphereRadius;
        d.dyn.grid.cs = make_float3(rp.particleRadius, rp.particleRadius, rp.particleRadius);
        d.dyn.grid.nb = (unsigned)rp.getDnNumBuckets();
        d.dyn.grid.mode = FG_KEY_REF_XOR;
        d.dyn.grid.key_base = (unsigned)nbB;
        nbB += (unsigned)rp.getDnNumBuckets();
        d.dyn.relation = make_int3(cp.fluid_to_srb_relation.x, cp.fluid_to_srb_relation.y, cp.fluid_to_srb_relation.z);
        d.dyn.fluid_radius = 2.0f * cp.pointRadius;
        d.dyn.body_radius = r;
        d.dyn.ks = -in.springForce;
        d.dyn.kd = -in.dampingForce;
        d.dyn.eps = 0.003f;
        d.dyn.enabled = 1;
        d.fluid.vmin = make_float3(cp.volume_min.x, cp.volume_min.y, cp.volume_min.z);
        d.fluid.cs = make_float3(cp.cell_size.x, cp.cell_size.y, cp.cell_size.z);
        d.fluid.dims = make_int3(cp.bucket_size.x, cp.bucket_size.y, cp.bucket_size.z);
        d.fluid.nb = (unsigned)cp.bucket_size.x * cp.bucket_size.y * cp.bucket_size.z;
        d.fluid.mode = FG_KEY_LINEAR;
        d.fluid.key_base = fd.keyBase;
        d.rb_to_fluid = make_int3(rp.rb_to_fluid_relation.x, rp.rb_to_fluid_relation.y, rp.rb_to_fluid_relation.z);
        d.sphere_radius = r;
        d.fluid_radius = 2.0f * cp.pointRadius;
        d.ks = -in.springForce;
        d.kd = -in.dampingForce;
        const float margin = 2.0f * cp.pointRadius + r + 0.003f;
        d.bounds_margin = margin * 1.01f;
        d.pair_margin = (2.0f * r + 0.001f) * 1.01f;
        d.dt = rp.deltaTimeIntegration;
        d.vlimit = rp.velocity_limit;
        d.ext = make_float4(rp.global_external_force.x, rp.global_external_force.y, rp.global_external_force.z,
                            rp.global_external_force.w);
        d.vmin = make_float3(rp.volume_min.x, rp.volume_min.y, rp.volume_min.z);
        d.vmax = make_float3(rp.volume_max.x, rp.volume_max.y, rp.volume_max.z);
    }
    const unsigned B = (unsigned)mass.size(), M = (unsigned)world.size();
    if (B == 0) return;
    if (nbB >= 0xFFFFFFFFull) throw std::invalid_argument("BatchedSimulator: too many body buckets");
    mBodyCount = B;
    mSphereCount = M;
    mBodyBuckets = (unsigned)nbB;

    d_bodyScene = dalloc<unsigned>(B);
    d_boffsets = dalloc<unsigned>(B + 1);
    d_bkey = dalloc<unsigned>(M);
    d_bid = dalloc<unsigned>(M);
    d_bstart = dalloc<unsigned>(mBodyBuckets);
    d_bend = dalloc<unsigned>(mBodyBuckets);
    d_brel = dalloc<float4>(M);
    d_bworld = dalloc<float4>(M);
    d_bsorted = dalloc<float4>(M);
    d_bpf = dalloc<float4>(M);
    d_bpt = dalloc<float4>(M);
    d_bcm = dalloc<float4>(B);
    d_bvel = dalloc<float4>(B);
    d_borient = dalloc<float4>(B);
    d_bL = dalloc<float4>(B);
    d_bforce = dalloc<float4>(B);
    d_btorque = dalloc<float4>(B);
    d_binv = dalloc<float4>(3 * B);
    d_sceneBounds = dalloc<float4>(2 * S);
    d_bodyBounds = dalloc<float4>(2 * B);
    d_defAcc = dalloc<float4>(mN);
    d_defList = dalloc<unsigned>(mN);
    d_defCount = dalloc<unsigned>(64);
    d_bmass = dalloc<float>(B);
    d_bsumF = dalloc<double>(4 * B);
    d_bsumT = dalloc<double>(4 * B);
    mBSortWsBytes = fg_sort_workspace_size(M, key_bits_for(mBodyBuckets));
    hip_check(hipMalloc(&d_bsortWs, mBSortWsBytes ? mBSortWsBytes : 256), "hipMalloc");
    hip_check(hipMalloc(&d_bscenes, sizeof(fg::BatchBodyScene) * S), "hipMalloc");

    hip_check(hipMemcpy(d_bworld, world.data(), sizeof(float4) * M, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_boffsets, offsets.data(), sizeof(unsigned) * (B + 1), hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_bodyScene, bodyScene.data(), sizeof(unsigned) * B, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_bmass, mass.data(), sizeof(float) * B, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemset(d_bend, 0, sizeof(unsigned) * mBodyBuckets), "memset");

    // Centre of mass and inertia in FP64 on the GPU (per body, as in the
    // single-fluid set-up), then the same host arithmetic.
    std::vector<double> md(mass.begin(), mass.end()), cmd(3 * B), inertia(9 * B);
    double *d_md = dalloc<double>(B), *d_cmd = dalloc<double>(3 * B), *d_in = dalloc<double>(9 * B);
    hip_check(hipMemcpy(d_md, md.data(), sizeof(double) * B, hipMemcpyHostToDevice), "upload");
    {
        StreamScope scope(mStream);
        fg_body_mass_properties_f64(reinterpret_cast<fg_float4*>(d_bworld), d_boffsets, B, d_md, d_cmd, d_in);
    }
    fg_check("fg_body_mass_properties_f64");
    hip_check(hipStreamSynchronize(mStream), "sync");
    hip_check(hipMemcpy(cmd.data(), d_cmd, sizeof(double) * 3 * B, hipMemcpyDeviceToHost), "download");
    hip_check(hipMemcpy(inertia.data(), d_in, sizeof(double) * 9 * B, hipMemcpyDeviceToHost), "download");
    dfree(d_md);
    dfree(d_cmd);
    dfree(d_in);
    std::vector<float4> cm(B), inv(3 * B), zero(B, make_float4(0, 0, 0, 0)), orient(B, make_float4(0, 0, 0, 1));
    for (unsigned b = 0; b < B; b++) {
        cm[b] = make_float4((float)cmd[3 * b], (float)cmd[3 * b + 1], (float)cmd[3 * b + 2], 0.0f);
        double iv[9];
        fg_host::invert3(&inertia[9 * b], iv);
        for (int r = 0; r < 3; r++)
            inv[3 * b + r] = make_float4((float)iv[3 * r], (float)iv[3 * r + 1], (float)iv[3 * r + 2], 0.0f);
    }
    std::vector<float4> rel(M);
    for (unsigned b = 0; b < B; b++)
        for (unsigned i = offsets[b]; i < offsets[b + 1]; i++)
            rel[i] = make_float4((float)((double)world[i].x - cmd[3 * b]), (float)((double)world[i].y - cmd[3 * b + 1]),
                                 (float)((double)world[i].z - cmd[3 * b + 2]), (float)b);
    hip_check(hipMemcpy(d_brel, rel.data(), sizeof(float4) * M, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_bcm, cm.data(), sizeof(float4) * B, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_binv, inv.data(), sizeof(float4) * 3 * B, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(d_borient, orient.data(), sizeof(float4) * B, hipMemcpyHostToDevice), "upload");
    for (float4* p : {d_bvel, d_bL, d_bforce, d_btorque})
        hip_check(hipMemcpy(p, zero.data(), sizeof(float4) * B, hipMemcpyHostToDevice), "upload");
    // the shared arrays each scene's contact walks
    for (unsigned s = 0; s < S; s++) {
        fg::BatchBodyScene& d = bs[s];
        d.dyn.pos = d_bsorted;
        d.dyn.vel = d_bvel;
        d.dyn.start = d_bstart;
        d.dyn.end = d_bend;
        d.dyn.aabb = d_sceneBounds + 2 * s;
        d.dyn.count = M;
        d.dyn.nvel = B;
    }
    hip_check(hipMemcpy(d_bscenes, bs.data(), sizeof(fg::BatchBodyScene) * S, hipMemcpyHostToDevice), "upload");
    mBodiesReady = true;
}

void BatchedSimulator::downloadBodies(float4* pos, float4* vel, float4* orient, float4* L) const {
    const size_t B = mBodyCount;
    if (!B) return;
    hip_check(hipStreamSynchronize(mStream), "sync");
    if (pos) hip_check(hipMemcpy(pos, d_bcm, sizeof(float4) * B, hipMemcpyDeviceToHost), "download");
    if (vel) hip_check(hipMemcpy(vel, d_bvel, sizeof(float4) * B, hipMemcpyDeviceToHost), "download");
    if (orient) hip_check(hipMemcpy(orient, d_borient, sizeof(float4) * B, hipMemcpyDeviceToHost), "download");
    if (L) hip_check(hipMemcpy(L, d_bL, sizeof(float4) * B, hipMemcpyDeviceToHost), "download");
}

void BatchedSimulator::invalidateGraph() {
    if (mGraphExec) (void)hipGraphExecDestroy(mGraphExec);
    if (mGraph) (void)hipGraphDestroy(mGraph);
    mGraphExec = nullptr;
    mGraph = nullptr;
}

void BatchedSimulator::setUseGraph(bool on) { mUseGraph = on; }
void BatchedSimulator::setOverlap(bool on) {
    if (on != mOverlap) invalidateGraph();
    mOverlap = on;
}
void BatchedSimulator::setStageTiming(bool on) { mTiming = on; }

void BatchedSimulator::setTuning(const BatchTuning& t) {
    if (t.queueItems < 0 || t.queueItems > 64)
        throw std::invalid_argument("BatchedSimulator::setTuning: queueItems outside 0..64");
    if (fg_set_tuning(t.sphBlock, t.sortDigitBits) != FG_OK) fg_check("setTuning");
    fg_set_schedule(t.flatDensity ? 1 : 0, t.flatForces ? 1 : 0);
    mTuning = t;
    invalidateGraph();  // the launch shapes are recorded in the graph
}

// Coupled step with the body stages on the body stream, in the dependency
// order of HeterogeneousSimulator::enqueueCoupledOverlapped:
//   bodies:  world transform, body grid, bounds   (last step's body state)
//   fluid:   hash, sort, cells -> [cells]; density
//   bodies:  forces from the sorted fluid         (after [cells])
//   fluid:   coupled forces + integration         (after the body grid)
//   bodies:  FP64 sums, integration, clears       (after the fluid forces,
//            which read the body velocities)
// and the fluid stream waits for the body stream at the end, so a step (and
// a graph captured from it) is complete when the fluid stream is.  Every
// kernel reads what it reads in the serial order.
void BatchedSimulator::enqueueStepOverlapped() {
    enum { kFork, kBodyGrid, kCells, kForces, kJoin };
    const unsigned S = (unsigned)mDescs.size();
    hip_check(hipEventRecord(mSync[kFork], mStream), "event");
    hip_check(hipStreamWaitEvent(mBodyStream, mSync[kFork], 0), "wait");
    fg_set_stream(mBodyStream);
    enqueueBodyStages(true);
    hip_check(hipEventRecord(mSync[kBodyGrid], mBodyStream), "event");
    fg_set_stream(mStream);
    fg_batch_hash(d_descs, S, d_sceneOf, reinterpret_cast<fg_float4*>(d_pos), d_key, mN);
    fg_sort_pairs_ws(d_key, d_id, mN, mKeyBits, d_sortWs, mSortWsBytes, 1);
    gFindCellOffsetOrderingPosVel(d_start, d_end, d_key, d_id, reinterpret_cast<fg_float4*>(d_pos),
                                  reinterpret_cast<fg_float4*>(d_spos), reinterpret_cast<fg_float4*>(d_vel),
                                  reinterpret_cast<fg_float4*>(d_svel), mN, mBuckets);
    hip_check(hipEventRecord(mSync[kCells], mStream), "event");
    fg::batch_density(d_descs, S, d_sceneOf, d_spos, d_svel, d_rho, d_start, d_end, d_masks, mN, work_queue(d_queue, mTuning));
    hip_check(hipStreamWaitEvent(mBodyStream, mSync[kCells], 0), "wait");
    fg_set_stream(mBodyStream);
    fg::batch_body_forces(d_bworld, d_bvel, d_bcm, d_bpf, d_bpt, mSphereCount, mBodyCount, d_bodyScene,
                          static_cast<const fg::BatchBodyScene*>(d_bscenes), d_spos, d_svel, d_start, d_end, mN,
                          d_bodyBounds);
    fg_set_stream(mStream);
    hip_check(hipStreamWaitEvent(mStream, mSync[kBodyGrid], 0), "wait");
    enqueueCoupledForces();
    hip_check(hipEventRecord(mSync[kForces], mStream), "event");
    hip_check(hipStreamWaitEvent(mBodyStream, mSync[kForces], 0), "wait");
    fg_set_stream(mBodyStream);
    enqueueBodyStages(false);
    hip_check(hipEventRecord(mSync[kJoin], mBodyStream), "event");
    fg_set_stream(mStream);
    hip_check(hipStreamWaitEvent(mStream, mSync[kJoin], 0), "wait");
}

// The coupled fluid force pass (forces + contact + walls + integration).
void BatchedSimulator::enqueueCoupledForces() {
    const unsigned S = (unsigned)mDescs.size();
    const fg::BatchBodyScene* bs = static_cast<const fg::BatchBodyScene*>(d_bscenes);
    if (mTuning.deferContact)
        fg::batch_forces_deferred(d_descs, S, d_sceneOf, d_spos, d_svel, d_rho, d_start, d_end, d_masks, d_vel, d_pos,
                                  d_id, mN, bs, fg::DeferredContact{d_defAcc, d_defList, d_defCount},
                                  work_queue(d_queue, mTuning));
    else
        fg::batch_forces_coupled(d_descs, S, d_sceneOf, d_spos, d_svel, d_rho, d_start, d_end, d_masks, d_vel, d_pos,
                                 d_id, mN, bs, work_queue(d_queue, mTuning));
}

// Stage order of HeterogeneousSimulator::doTimeStep for the fused fast path.
void BatchedSimulator::enqueueStep() {
    if (mBodyCount && mOverlap && !mTiming) {
        enqueueStepOverlapped();
        return;
    }
    const unsigned S = (unsigned)mDescs.size();
    if (mTiming) (void)hipEventRecord(mEv[0], mStream);
    fg_batch_hash(d_descs, S, d_sceneOf, reinterpret_cast<fg_float4*>(d_pos), d_key, mN);
    if (mTiming) (void)hipEventRecord(mEv[1], mStream);
    fg_sort_pairs_ws(d_key, d_id, mN, mKeyBits, d_sortWs, mSortWsBytes, 1);
    if (mTiming) (void)hipEventRecord(mEv[2], mStream);
    gFindCellOffsetOrderingPosVel(d_start, d_end, d_key, d_id, reinterpret_cast<fg_float4*>(d_pos),
                                  reinterpret_cast<fg_float4*>(d_spos), reinterpret_cast<fg_float4*>(d_vel),
                                  reinterpret_cast<fg_float4*>(d_svel), mN, mBuckets);
    if (mBodyCount) enqueueBodyStages(true);
    if (mTiming) (void)hipEventRecord(mEv[3], mStream);
    fg::batch_density(d_descs, S, d_sceneOf, d_spos, d_svel, d_rho, d_start, d_end, d_masks, mN, work_queue(d_queue, mTuning));
    if (mTiming) (void)hipEventRecord(mEv[4], mStream);
    if (mBodyCount) {
        fg::batch_body_forces(d_bworld, d_bvel, d_bcm, d_bpf, d_bpt, mSphereCount, mBodyCount, d_bodyScene,
                              static_cast<const fg::BatchBodyScene*>(d_bscenes), d_spos, d_svel, d_start, d_end, mN,
                              d_bodyBounds);
        enqueueCoupledForces();
        enqueueBodyStages(false);
    } else {
        fg::batch_forces(d_descs, S, d_sceneOf, d_spos, d_svel, d_rho, d_start, d_end, d_masks, d_vel, d_pos, d_id,
                         mN, work_queue(d_queue, mTuning));
    }
    if (mTiming) (void)hipEventRecord(mEv[5], mStream);
}

// Body stages of the coupled step (RigidBodySimulator::generateCollisionData
// before the fluid density; applyForces and integrate after the fluid
// forces), for every scene at once.
void BatchedSimulator::enqueueBodyStages(bool before_density) {
    const fg::BatchBodyScene* bs = static_cast<const fg::BatchBodyScene*>(d_bscenes);
    if (before_density) {
        fg::batch_body_transform(d_bkey, d_bid, d_bcm, d_borient, d_brel, d_bworld, mSphereCount, d_bodyScene, bs);
        fg_sort_pairs_ws(d_bkey, d_bid, mSphereCount, key_bits_for(mBodyBuckets), d_bsortWs, mBSortWsBytes, 1);
        gFindCellOffsetOrderingPos(d_bstart, d_bend, d_bkey, d_bid, reinterpret_cast<fg_float4*>(d_bworld),
                                   reinterpret_cast<fg_float4*>(d_bsorted), mSphereCount, mBodyBuckets);
        fg::batch_scene_bounds(d_bworld, bs, (unsigned)mDescs.size(), d_sceneBounds);
        fg::batch_body_bounds(d_bworld, d_boffsets, d_bodyScene, bs, mBodyCount, d_bodyBounds);
        return;
    }
    fg_body_reduce_f64(reinterpret_cast<fg_float4*>(d_bpf), reinterpret_cast<fg_float4*>(d_bpt), d_boffsets,
                       mBodyCount, reinterpret_cast<fg_double4*>(d_bsumF), reinterpret_cast<fg_double4*>(d_bsumT),
                       reinterpret_cast<fg_float4*>(d_bforce), reinterpret_cast<fg_float4*>(d_btorque));
    fg::batch_integrate_bodies(d_bcm, d_bmass, d_bvel, d_bforce, mBodyCount, d_bodyScene, bs, d_bL, d_btorque, d_binv,
                               d_borient);
    fg_memset_u32(reinterpret_cast<unsigned*>(d_bforce), 0u, 4 * (size_t)mBodyCount);
    fg_memset_u32(reinterpret_cast<unsigned*>(d_btorque), 0u, 4 * (size_t)mBodyCount);
}

void BatchedSimulator::doTimeStep() {
    setupBodies();
    StreamScope scope(mStream);
    enqueueStep();
    fg_check("BatchedSimulator::doTimeStep");
    if (mTiming) {
        hip_check(hipEventSynchronize(mEv[5]), "stage timing");
        for (int k = 0; k < 5; k++) (void)hipEventElapsedTime(&mTimes[k], mEv[k], mEv[k + 1]);
    }
}

void BatchedSimulator::doTimeSteps(int steps) {
    if (steps <= 0) return;
    setupBodies();
    if (!mUseGraph || mTiming) {
        for (int s = 0; s < steps; s++) doTimeStep();
        return;
    }
    StreamScope scope(mStream);
    if (!mGraphExec) {
        hip_check(hipStreamBeginCapture(mStream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        enqueueStep();
        hipGraph_t g = nullptr;
        hip_check(hipStreamEndCapture(mStream, &g), "hipStreamEndCapture");
        fg_check("capture");
        mGraph = g;
        hip_check(hipGraphInstantiate(&mGraphExec, mGraph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    for (int s = 0; s < steps; s++) hip_check(hipGraphLaunch(mGraphExec, mStream), "hipGraphLaunch");
}

void BatchedSimulator::synchronize() { hip_check(hipStreamSynchronize(mStream), "hipStreamSynchronize"); }

BatchTuning BatchedSimulator::autotune(int trialSteps, std::vector<std::pair<BatchTuning, float>>* trials) {
    if (trialSteps < 1) trialSteps = 1;
    setupBodies();
    synchronize();
    // Save the state so that the trials leave the simulation where it was.
    float4 *bp = dalloc<float4>(mN), *bv = dalloc<float4>(mN);
    hip_check(hipMemcpy(bp, d_pos, sizeof(float4) * mN, hipMemcpyDeviceToDevice), "save");
    hip_check(hipMemcpy(bv, d_vel, sizeof(float4) * mN, hipMemcpyDeviceToDevice), "save");
    const size_t B = mBodyCount;
    float4* const bodyState[4] = {d_bcm, d_bvel, d_borient, d_bL};
    float4* bb = B ? dalloc<float4>(4 * B) : nullptr;
    for (int k = 0; k < 4 && B; k++)
        hip_check(hipMemcpy(bb + k * B, bodyState[k], sizeof(float4) * B, hipMemcpyDeviceToDevice), "save");
    auto restore_bodies = [&]() {
        for (int k = 0; k < 4 && B; k++)
            hip_check(hipMemcpyAsync(bodyState[k], bb + k * B, sizeof(float4) * B, hipMemcpyDeviceToDevice, mStream),
                      "restore");
    };
    const bool timing = mTiming, graph = mUseGraph;
    mTiming = false;
    mUseGraph = true;
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    // Every trial starts from the saved state, so all shapes are timed over
    // the same particle configurations.
    auto time_shape = [&](const BatchTuning& t, bool record = true) {
        setTuning(t);
        hip_check(hipMemcpyAsync(d_pos, bp, sizeof(float4) * mN, hipMemcpyDeviceToDevice, mStream), "restore");
        hip_check(hipMemcpyAsync(d_vel, bv, sizeof(float4) * mN, hipMemcpyDeviceToDevice, mStream), "restore");
        restore_bodies();
        doTimeSteps(2);  // capture + warm-up
        hip_check(hipEventRecord(e0, mStream), "event");
        doTimeSteps(trialSteps);
        hip_check(hipEventRecord(e1, mStream), "event");
        hip_check(hipEventSynchronize(e1), "event");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
        ms /= (float)trialSteps;
        if (trials && record) trials->push_back({t, ms});
        return ms;
    };
    // one discarded trial first: the clocks of an idle GPU ramp up over it
    (void)time_shape(mTuning, false);
    // Coordinate descent from the current shape: workgroup size, schedules,
    // work queue, contact, digit width.  A candidate replaces the incumbent
    // only when it is faster by more than kTuneMargin, about the spread of
    // repeated trials of one shape.
    constexpr float kTuneMargin = 0.01f;
    BatchTuning best = mTuning;
    float best_ms = time_shape(best);
    auto better = [&](float ms) { return ms < best_ms * (1.0f - kTuneMargin); };
    for (int blk : {64, 128, 256, 512}) {
        BatchTuning t = best;
        t.sphBlock = blk;
        if (t.sphBlock == best.sphBlock) continue;
        const float ms = time_shape(t);
        if (better(ms)) { best_ms = ms; best = t; }
    }
    for (int sched = 0; sched < 4; sched++) {
        BatchTuning t = best;
        t.flatDensity = (sched & 1) != 0;
        t.flatForces = (sched & 2) != 0;
        if (t.flatDensity == best.flatDensity && t.flatForces == best.flatForces) continue;
        const float ms = time_shape(t);
        if (better(ms)) { best_ms = ms; best = t; }
    }
    for (int items : {0, 1, 4}) {
        BatchTuning t = best;
        t.queueItems = items;
        if (t.queueItems == best.queueItems) continue;
        const float ms = time_shape(t);
        if (better(ms)) { best_ms = ms; best = t; }
    }
    if (mBodyCount) {
        BatchTuning t = best;
        t.deferContact = !best.deferContact;
        const float ms = time_shape(t);
        if (better(ms)) { best_ms = ms; best = t; }
    }
    for (int bits : {8, 7, 6}) {
        BatchTuning t = best;
        t.sortDigitBits = bits;
        if (t.sortDigitBits == best.sortDigitBits && !trials) continue;
        const float ms = time_shape(t);
        if (better(ms)) { best_ms = ms; best = t; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    setTuning(best);
    hip_check(hipMemcpyAsync(d_pos, bp, sizeof(float4) * mN, hipMemcpyDeviceToDevice, mStream), "restore");
    hip_check(hipMemcpyAsync(d_vel, bv, sizeof(float4) * mN, hipMemcpyDeviceToDevice, mStream), "restore");
    restore_bodies();
    hip_check(hipStreamSynchronize(mStream), "restore");
    dfree(bp);
    dfree(bv);
    if (bb) {
        hip_check(hipStreamSynchronize(mStream), "sync");
        dfree(bb);
    }
    mTiming = timing;
    mUseGraph = graph;
    return best;
}

void BatchedSimulator::download(float4* positions, float4* velocities) const {
    hip_check(hipStreamSynchronize(mStream), "sync");
    if (positions) hip_check(hipMemcpy(positions, d_pos, sizeof(float4) * mN, hipMemcpyDeviceToHost), "download");
    if (velocities) hip_check(hipMemcpy(velocities, d_vel, sizeof(float4) * mN, hipMemcpyDeviceToHost), "download");
}

void BatchedSimulator::downloadDensity(float2* rho_p_sorted, unsigned* particleId) const {
    hip_check(hipStreamSynchronize(mStream), "sync");
    if (rho_p_sorted) hip_check(hipMemcpy(rho_p_sorted, d_rho, sizeof(float2) * mN, hipMemcpyDeviceToHost), "download");
    if (particleId) hip_check(hipMemcpy(particleId, d_id, sizeof(unsigned) * mN, hipMemcpyDeviceToHost), "download");
}

}  // namespace Simulators
}  // namespace Simulation
