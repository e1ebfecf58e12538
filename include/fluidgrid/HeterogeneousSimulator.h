// HeterogeneousSimulator.h -- the step orchestrator of the reference
// (HeterogeneousSimulator.h:41-45, .cpp:27-212): a singleton that computes
// the kernel constants, uploads the parameter blocks and advances every
// registered fluid by one step per doTimeStep().
//
// Differences from the reference, all outside the arithmetic:
//  * no host synchronisation or printf between stages; work goes to one
//    stream (the simulator's own, or the one given to setStream);
//  * the sort uses the key width implied by the bucket count and a
//    persistent workspace; forces and integration run as one pass when
//    `fused` is set (bit-identical results);
//  * doTimeSteps(n) can replay a captured HIP graph of one step;
//  * with rigid bodies, the body work runs on a second stream beside the
//    fluid's hash, sort and density (same dependencies, same results).
#pragma once

#include <string>
#include <vector>

#include "ApplicationData.h"
#include "GPUSPHFluid.h"
#include "RigidBodySimulator.h"

namespace Simulation {
namespace Simulators {

struct StepOptions {
    bool fused = true;       // forces + integrate in one pass
    bool use_graph = true;   // doTimeSteps replays a captured step
    bool stage_timing = false;
    bool overlap_bodies = true;  // body work on a second stream (results unchanged)
    // Density arithmetic of the fused step (FG_PRECISION_FP32 or
    // FG_PRECISION_MIXED); the unfused step is always FP32.
    int precision = FG_PRECISION_FP32;
};

// Launch shape of the single-fluid passes (process-wide library state:
// fg_set_tuning, fg_set_schedule, fg_set_tiled).  No choice changes any
// result bit.
struct StepTuning {
    int sphBlock = 256;        // workgroup size of the fast density/force kernels
    int sortDigitBits = 8;     // widest radix digit
    bool flatDensity = false;  // density candidates as one stream per lane
    bool flatForces = false;   // force neighbours as one stream per lane
    bool tiled = false;        // LDS-tiled density and force passes
};

struct StageTimes {
    float hash_ms = 0, sort_ms = 0, cells_ms = 0, density_ms = 0, forces_ms = 0, integrate_ms = 0, bodies_ms = 0;
};

class HeterogeneousSimulator {
public:
    static HeterogeneousSimulator* instance();
    // Extension: drops the singleton (and its captured graph).
    static void destroy();

    void addSphFluid(Fluid::GPUSPHFluid* fluid);
    void doTimeStep();

    // Extensions.
    void doTimeSteps(int steps);
    void setOptions(const StepOptions& o);
    const StepOptions& options() const { return mOptions; }
    void setStream(hipStream_t s);
    hipStream_t stream() const { return mStream; }
    const StageTimes& lastStageTimes() const { return mTimes; }
    void invalidateGraph();
    // Checkpoint / resume (no counterpart in the reference).  saveState
    // writes everything a step reads from the one before: per fluid the
    // positions and velocities (particle order), the sorted (rho, p) buffer
    // and the inflow step counter; per dynamic body the centre of mass,
    // velocity, orientation and angular momentum.  The parameter blocks go
    // along to check the scene.  loadState restores such a file into a
    // simulator set up with the same scene (same parameter blocks, fluids,
    // particle counts, inflows and bodies); continuing from it gives the
    // same bits as not having stopped.  Both throw std::runtime_error on a
    // mismatch, a damaged file (checksum) or an I/O error; loadState
    // changes nothing unless the whole file checks.
    void saveState(const std::string& path);
    void loadState(const std::string& path);
    // Times launch shapes by coordinate descent (workgroup size, then the
    // density and force schedules, then the LDS-tiled passes, then the
    // digit width; trialSteps graph steps each after a capture and a warm-up
    // step, every trial from the same state) and keeps the fastest; the
    // simulation state (particles, (rho, p), inflow counters, bodies) is
    // restored afterwards, so tuning does not advance the simulation.
    StepTuning autotune(int trialSteps, std::vector<std::pair<StepTuning, float>>* trials = nullptr);
    static void applyTuning(const StepTuning& t);
    static StepTuning currentTuning();

private:
    HeterogeneousSimulator();
    ~HeterogeneousSimulator();

    void findFluidNeighbourhood(Fluid::GPUSPHFluid* fluid);
    void calculateFluidDensity(Fluid::GPUSPHFluid* fluid);
    void calculateFluidForces(Fluid::GPUSPHFluid* fluid, bool useStaticRB, bool useDynamicRB);
    void integrateFluidForces(Fluid::GPUSPHFluid* fluid);
    void enqueueStep();
    void enqueueCoupledOverlapped(Fluid::GPUSPHFluid* fluid);

    std::vector<Fluid::GPUSPHFluid*> mFluids;
    StepOptions mOptions;
    hipStream_t mStream = nullptr;
    bool mOwnStream = false;
    hipGraph_t mGraph = nullptr;
    hipGraphExec_t mGraphExec = nullptr;
    hipEvent_t mEv[9] = {};
    hipStream_t mBodyStream = nullptr;  // body work of a coupled step
    hipEvent_t mSync[5] = {};           // fork/join of the two streams
    void emit(Fluid::GPUSPHFluid* fluid);
    bool coupled() const;        // dynamic bodies present
    bool hasStaticBodies() const;
    StageTimes mTimes;
    static HeterogeneousSimulator* mInstance;
};

}  // namespace Simulators
}  // namespace Simulation
