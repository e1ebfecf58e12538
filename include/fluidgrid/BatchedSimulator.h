// BatchedSimulator.h -- many independent SPH fluids ("scenes") advanced
// together with one launch per stage.  No counterpart in the reference,
// whose HeterogeneousSimulator shares one CommonFluidParams between all
// fluids (HeterogeneousSimulator.cpp:27-47); here every scene keeps its own
// parameter block, grid and kernel constants.
//
// Per step: keys (scene key base + LINEAR key) -> one stable sort of the
// shared keys -> cell ranges + reorder over the total bucket count ->
// density (+ neighbour masks) -> forces + walls + integration.  The step has
// no host synchronisation and is replayed as a captured HIP graph.
#pragma once

#include <stdexcept>
#include <vector>

#include "ApplicationData.h"

namespace Simulation {
namespace Simulators {

// A scene the batched passes cannot step correctly (cells smaller than the
// kernel radius); fg_bsim_create reports it as FG_ERR_UNSUPPORTED.
struct UnsupportedScene : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

struct BatchScene {
    Application::CommonFluidParams common;  // LINEAR grid of this scene
    Application::FluidParams fluid;
    std::vector<float4> positions;          // w = active flag
    std::vector<float4> velocities;
};

// Dynamic rigid bodies of one scene: the inputs a single-fluid simulator
// takes for them (fg_sim_add_bodies).
struct BatchBodies {
    std::vector<float4> worldSpheres;  // body b owns [offsets[b], offsets[b+1]); w is ignored
    std::vector<unsigned> offsets;     // numBodies + 1 entries, offsets[0] = 0
    std::vector<double> masses;        // kg per body
    float sphereRadius = 0.0f;
    Application::CommonRigidBodyParams rigid{};
    float springForce = 0.0f, dampingForce = 0.0f;
};

struct BatchTuning {
    int sphBlock = 256;       // workgroup size of density and force kernels
    int sortDigitBits = 8;    // widest radix digit
    bool flatDensity = false; // density candidates as one stream per lane
    bool flatForces = false;  // force neighbours as one stream per lane
    // Persistent density and force launches taking work items of this many
    // workgroups' particles from a queue (0: one workgroup per block).
    int queueItems = 0;
    // Coupled force pass in two passes: particles outside every body grid's
    // bounds are finished by a force kernel without the contact walk, the
    // others by a second pass with it (same bits).
    bool deferContact = true;
};

class BatchedSimulator {
public:
    explicit BatchedSimulator(const std::vector<BatchScene>& scenes);
    ~BatchedSimulator();
    BatchedSimulator(const BatchedSimulator&) = delete;
    BatchedSimulator& operator=(const BatchedSimulator&) = delete;

    void doTimeStep();                 // eager; records stage events when timing is on
    void doTimeSteps(int steps);       // graph replay (eager when timing is on)
    void setUseGraph(bool on);
    void setStageTiming(bool on);
    // Body stages on a second stream beside the fluid stages (default on;
    // results unchanged; stage timing runs them in line).
    void setOverlap(bool on);
    void synchronize();

    // Times launch shapes and schedules by coordinate descent (workgroup
    // size, then the density and force schedules, then the persistent
    // work-queue launches (none, 1 or 4 workgroups per item), then the
    // deferred contact pass (with bodies), then the digit width;
    // trialSteps graph steps each after a warm-up, every trial from the
    // current particle state) and keeps the fastest; the state is restored
    // afterwards, so tuning does not advance the simulation.  No choice
    // changes any result bit.
    BatchTuning autotune(int trialSteps, std::vector<std::pair<BatchTuning, float>>* trials = nullptr);
    void setTuning(const BatchTuning& t);
    const BatchTuning& tuning() const { return mTuning; }

    void download(float4* positions, float4* velocities) const;
    void downloadDensity(float2* rho_p_sorted, unsigned* particleId) const;
    // hash, sort, cells, density, forces+integrate (ms) of the last timed step
    const float* stageTimes() const { return mTimes; }

    // Dynamic rigid bodies of scene s, coupled two-way with that scene's
    // fluid with the stages and arithmetic of the single-fluid coupled step
    // (HeterogeneousSimulator with RigidBodySimulator); bodies of different
    // scenes never interact.  Before the first step, at most once per scene.
    void addBodies(unsigned scene, const BatchBodies& bodies);
    unsigned numBodies() const { return mBodyCount; }
    // Per body, scenes in order: centre of mass, velocity, orientation
    // (x, y, z, w), angular momentum.  Any pointer may be null.
    void downloadBodies(float4* pos, float4* vel, float4* orient, float4* L) const;

    unsigned numScenes() const { return (unsigned)mDescs.size(); }
    unsigned numParticles() const { return mN; }
    unsigned numBuckets() const { return mBuckets; }
    int keyBits() const { return mKeyBits; }
    hipStream_t stream() const { return mStream; }

private:
    void enqueueStep();
    void enqueueStepOverlapped();
    void enqueueCoupledForces();
    void enqueueBodyStages(bool before_density);
    void invalidateGraph();
    void setupBodies();
    void releaseBodies();

    std::vector<fg_scene_desc> mDescs;
    unsigned mN = 0, mBuckets = 0;
    int mKeyBits = 1;
    BatchTuning mTuning;
    bool mUseGraph = true, mTiming = false, mOverlap = true;
    hipStream_t mStream = nullptr;
    hipStream_t mBodyStream = nullptr;
    hipEvent_t mSync[5] = {};
    hipGraph_t mGraph = nullptr;
    hipGraphExec_t mGraphExec = nullptr;
    hipEvent_t mEv[6] = {};
    float mTimes[5] = {};

    fg_scene_desc* d_descs = nullptr;
    unsigned* d_sceneOf = nullptr;
    float4 *d_pos = nullptr, *d_vel = nullptr, *d_spos = nullptr, *d_svel = nullptr;
    float2* d_rho = nullptr;
    unsigned *d_key = nullptr, *d_id = nullptr, *d_start = nullptr, *d_end = nullptr;
    unsigned long long* d_masks = nullptr;
    unsigned* d_queue = nullptr;  // work-queue counters of the persistent launches
    void* d_sortWs = nullptr;
    size_t mSortWsBytes = 0;

    // rigid bodies (all scenes; bodies scene-major, spheres in body order)
    std::vector<BatchBodies> mSceneBodies;
    bool mBodiesReady = false;
    unsigned mBodyCount = 0, mSphereCount = 0, mBodyBuckets = 0;
    void* d_bscenes = nullptr;  // per-scene body descriptors (device)
    unsigned *d_bodyScene = nullptr, *d_boffsets = nullptr, *d_bkey = nullptr, *d_bid = nullptr;
    unsigned *d_bstart = nullptr, *d_bend = nullptr;
    float4 *d_brel = nullptr, *d_bworld = nullptr, *d_bsorted = nullptr, *d_bpf = nullptr, *d_bpt = nullptr;
    float4 *d_bcm = nullptr, *d_bvel = nullptr, *d_borient = nullptr, *d_bL = nullptr, *d_bforce = nullptr;
    float4 *d_btorque = nullptr, *d_binv = nullptr, *d_sceneBounds = nullptr, *d_bodyBounds = nullptr;
    float* d_bmass = nullptr;
    double *d_bsumF = nullptr, *d_bsumT = nullptr;
    float4* d_defAcc = nullptr;  // deferred contact: pair sums, particle list and count
    unsigned *d_defList = nullptr, *d_defCount = nullptr;
    void* d_bsortWs = nullptr;
    size_t mBSortWsBytes = 0;
};

}  // namespace Simulators
}  // namespace Simulation
