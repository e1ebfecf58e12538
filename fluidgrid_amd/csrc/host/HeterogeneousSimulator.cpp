// HeterogeneousSimulator.cpp -- per-step orchestration of the SPH pipeline
// (HeterogeneousSimulator.cpp:19-212 of the reference).
#include "fluidgrid/HeterogeneousSimulator.h"

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace Simulation {
namespace Simulators {

HeterogeneousSimulator* HeterogeneousSimulator::mInstance = nullptr;

namespace {
void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void fg_check(const char* what) {
    int e = fg_peek_error();
    if (e != FG_OK) {
        std::string msg = std::string(what) + ": " + fg_last_error_string();
        (void)fg_last_error();
        throw std::runtime_error(msg);
    }
}

int key_bits_for(unsigned int numBuckets) {
    int b = 1;
    while (b < 32 && ((numBuckets - 1u) >> b) != 0) b++;
    return b;
}

// Makes `s` the launch stream for the duration of a scope.
struct StreamScope {
    void* prev;
    explicit StreamScope(hipStream_t s) : prev(fg_get_stream()) { fg_set_stream(s); }
    ~StreamScope() { fg_set_stream(prev); }
};

// Makes `mode` the density arithmetic for the duration of a scope.
struct PrecisionScope {
    int prev;
    explicit PrecisionScope(int mode) : prev(fg_get_precision()) { fg_set_precision(mode); }
    ~PrecisionScope() { fg_set_precision(prev); }
};

fg_float3 abi3(float3 v) { return fg_float3{v.x, v.y, v.z}; }
}  // namespace

HeterogeneousSimulator* HeterogeneousSimulator::instance() {
    if (mInstance == nullptr) mInstance = new HeterogeneousSimulator();
    return mInstance;
}

void HeterogeneousSimulator::destroy() {
    delete mInstance;
    mInstance = nullptr;
}

// HeterogeneousSimulator.cpp:27-47: kernel constants with the float PI and
// pow(h, n) in double, one rounding to float; then both parameter blocks
// are handed to the launchers.
HeterogeneousSimulator::HeterogeneousSimulator() {
    using namespace Application;
    CommonFluidParams& cp = ApplicationData::instance()->getCommonFluidParams();
    KernelsCache kc;
    const double h = cp.kernel_radius;
    kc.kPoly6 = (float)(315.0f / (64.0f * PI * std::pow(h, 9.0)));
    kc.kSpikyGradient = (float)(-(45.0f / (PI * std::pow(h, 6.0))));
    kc.kViscosityLaplacian = (float)(45.0f / (PI * std::pow(h, 6.0)));
    gInitializeKernelParams(abi_cast<fg_kernels_cache>(kc));
    ApplicationData::instance()->setKernelsCache(kc);
    gInitializeCommonFluidParams(abi_cast<fg_common_fluid_params>(cp));

    hip_check(hipStreamCreateWithFlags(&mStream, hipStreamNonBlocking), "hipStreamCreate");
    mOwnStream = true;
    hip_check(hipStreamCreateWithFlags(&mBodyStream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : mEv) hip_check(hipEventCreate(&e), "hipEventCreate");
    for (auto& e : mSync) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
}

HeterogeneousSimulator::~HeterogeneousSimulator() {
    invalidateGraph();
    if (mStream) (void)hipStreamSynchronize(mStream);
    if (mBodyStream) (void)hipStreamSynchronize(mBodyStream);
    for (auto& e : mEv)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : mSync)
        if (e) (void)hipEventDestroy(e);
    if (mBodyStream) (void)hipStreamDestroy(mBodyStream);
    if (mOwnStream && mStream) (void)hipStreamDestroy(mStream);
}

void HeterogeneousSimulator::addSphFluid(Fluid::GPUSPHFluid* fluid) {
    mFluids.push_back(fluid);
    invalidateGraph();
}

void HeterogeneousSimulator::setOptions(const StepOptions& o) {
    mOptions = o;
    invalidateGraph();
}

void HeterogeneousSimulator::setStream(hipStream_t s) {
    invalidateGraph();
    if (mOwnStream && mStream) {
        (void)hipStreamSynchronize(mStream);
        (void)hipStreamDestroy(mStream);
    }
    mStream = s;
    mOwnStream = false;
}

void HeterogeneousSimulator::invalidateGraph() {
    if (mGraphExec) (void)hipGraphExecDestroy(mGraphExec);
    if (mGraph) (void)hipGraphDestroy(mGraph);
    mGraphExec = nullptr;
    mGraph = nullptr;
}

// HeterogeneousSimulator.cpp:131-150
void HeterogeneousSimulator::findFluidNeighbourhood(Fluid::GPUSPHFluid* fluid) {
    const Application::CommonFluidParams& cp = Application::ApplicationData::instance()->getCommonFluidParams();
    const unsigned int nb = (unsigned int)cp.getNumBuckets();
    const unsigned int n = (unsigned int)fluid->getNumParticles();
    if (mOptions.fused) {
        // Keys only; the first sort pass generates the identity ids.  The
        // last step's sorted keys, which the new keys replace, name the only
        // cells whose start is set: clearing those leaves cellStart empty for
        // this step's cell pass, which then skips the fill of every bucket.
        fg_calc_hash_key_clear(fluid->getHashKey(), reinterpret_cast<fg_float4*>(fluid->getPosition()), n, nb,
                               abi3(cp.cell_size), fluid->getHashStart());
    } else {
        gCalcHashKey(fluid->getHashKey(), fluid->getParticleID(), reinterpret_cast<fg_float4*>(fluid->getPosition()),
                     n, nb, abi3(cp.cell_size));
    }
    if (mOptions.stage_timing) (void)hipEventRecord(mEv[1], mStream);
    fg_sort_pairs_ws(fluid->getHashKey(), fluid->getParticleID(), n, key_bits_for(nb), fluid->getSortWorkspace(),
                     fluid->getSortWorkspaceBytes(), mOptions.fused ? 1 : 0);
    if (mOptions.stage_timing) (void)hipEventRecord(mEv[2], mStream);
    if (mOptions.fused) {
        fg_cell_offsets_pos_vel(fluid->getHashStart(), fluid->getHashEnd(), fluid->getHashKey(),
                                fluid->getParticleID(), reinterpret_cast<fg_float4*>(fluid->getPosition()),
                                reinterpret_cast<fg_float4*>(fluid->getOrderedPosition()),
                                reinterpret_cast<fg_float4*>(fluid->getVelocity()),
                                reinterpret_cast<fg_float4*>(fluid->getOrderedVelocities()), n, nb, 0,
                                fg_get_precision() == FG_PRECISION_MIXED ? fluid->getNeighborMasks() : nullptr);
        return;
    }
    gFindCellOffsetOrderingPosVel(fluid->getHashStart(), fluid->getHashEnd(), fluid->getHashKey(),
                                  fluid->getParticleID(), reinterpret_cast<fg_float4*>(fluid->getPosition()),
                                  reinterpret_cast<fg_float4*>(fluid->getOrderedPosition()),
                                  reinterpret_cast<fg_float4*>(fluid->getVelocity()),
                                  reinterpret_cast<fg_float4*>(fluid->getOrderedVelocities()), n, nb);
}

// HeterogeneousSimulator.cpp:153-159
void HeterogeneousSimulator::calculateFluidDensity(Fluid::GPUSPHFluid* fluid) {
    if (mOptions.fused) {
        // Same density, plus the neighbour bitmasks the force pass reuses.
        fg_sph_density_nbr(reinterpret_cast<fg_float4*>(fluid->getOrderedPosition()),
                           reinterpret_cast<const fg_float4*>(fluid->getOrderedVelocities()),
                           reinterpret_cast<fg_float2*>(fluid->getDensityField()), fluid->getHashStart(),
                           fluid->getHashEnd(), fluid->getNeighborMasks(), (unsigned)fluid->getNumParticles(),
                           fluid->getGasConstant(), fluid->getRestDensity(), fluid->getMass());
        return;
    }
    gCalcMassDensity(reinterpret_cast<fg_float4*>(fluid->getOrderedPosition()),
                     reinterpret_cast<fg_float2*>(fluid->getDensityField()), fluid->getHashKey(),
                     fluid->getHashStart(), fluid->getHashEnd(), fluid->getNumParticles(), fluid->getGasConstant(),
                     fluid->getRestDensity(), fluid->getMass());
}

// Inflow emission before the neighbourhood search (the reference never calls
// IEmitter::update; its only emitter's update is empty, Emitter.h:46).
void HeterogeneousSimulator::emit(Fluid::GPUSPHFluid* f) {
    if (f->hasInflow())
        f->emitInflow(Application::ApplicationData::instance()->getCommonFluidParams().deltaTimeIntegration);
}

bool HeterogeneousSimulator::coupled() const {
    RigidBody::RigidBodySimulator* rb = RigidBody::RigidBodySimulator::instance();
    return rb->getNumDynamicBodies() > 0;
}

bool HeterogeneousSimulator::hasStaticBodies() const {
    return RigidBody::RigidBodySimulator::instance()->getNumStaticSpheres() > 0;
}

// HeterogeneousSimulator.cpp:161-205.  The coupled branch (commented out in
// the reference, :174-186) is the static and dynamic grids of the rigid-body
// simulator with the fluid contact radius 2*pointRadius (:179).
void HeterogeneousSimulator::calculateFluidForces(Fluid::GPUSPHFluid* fluid, bool useStaticRB, bool useDynamicRB) {
    if (useStaticRB || useDynamicRB) {
        const Application::CommonFluidParams& cp = Application::ApplicationData::instance()->getCommonFluidParams();
        RigidBody::RigidBodySimulator* rb = RigidBody::RigidBodySimulator::instance();
        const fg_body_grid g = rb->bodyGrid(2.0f * cp.pointRadius);
        const fg_body_grid st = rb->staticGrid(2.0f * cp.pointRadius);
        fg_sph_forces_coupled(reinterpret_cast<fg_float4*>(fluid->getOrderedPosition()),
                              reinterpret_cast<fg_float4*>(fluid->getOrderedVelocities()),
                              reinterpret_cast<fg_float2*>(fluid->getDensityField()), fluid->getHashStart(),
                              fluid->getHashEnd(), mOptions.fused ? fluid->getNeighborMasks() : nullptr,
                              reinterpret_cast<fg_float4*>(fluid->getVelocity()),
                              mOptions.fused ? reinterpret_cast<fg_float4*>(fluid->getPosition()) : nullptr,
                              fluid->getParticleID(), (unsigned)fluid->getNumParticles(), fluid->getMass(),
                              fluid->getViscosity(), useDynamicRB ? &g : nullptr, useStaticRB ? &st : nullptr);
        return;
    }
    if (mOptions.fused) {
        fg_sph_forces_integrate_nbr(reinterpret_cast<fg_float4*>(fluid->getOrderedPosition()),
                                    reinterpret_cast<fg_float4*>(fluid->getOrderedVelocities()),
                                    reinterpret_cast<fg_float2*>(fluid->getDensityField()), fluid->getHashStart(),
                                    fluid->getHashEnd(), fluid->getNeighborMasks(),
                                    reinterpret_cast<fg_float4*>(fluid->getVelocity()),
                                    reinterpret_cast<fg_float4*>(fluid->getPosition()), fluid->getParticleID(),
                                    (unsigned)fluid->getNumParticles(), fluid->getMass(), fluid->getViscosity());
    } else {
        gSPHCalculateForces(reinterpret_cast<fg_float4*>(fluid->getOrderedPosition()),
                            reinterpret_cast<fg_float4*>(fluid->getOrderedVelocities()),
                            reinterpret_cast<fg_float4*>(fluid->getForces()),
                            reinterpret_cast<fg_float2*>(fluid->getDensityField()), fluid->getHashStart(),
                            fluid->getHashEnd(), reinterpret_cast<fg_float4*>(fluid->getVelocity()),
                            fluid->getParticleID(), (unsigned)fluid->getNumParticles(), fluid->getMass(),
                            fluid->getViscosity());
    }
}

// HeterogeneousSimulator.cpp:207-212
void HeterogeneousSimulator::integrateFluidForces(Fluid::GPUSPHFluid* fluid) {
    gSPHIntegrate(reinterpret_cast<fg_float4*>(fluid->getPosition()),
                  reinterpret_cast<fg_float4*>(fluid->getVelocity()), reinterpret_cast<fg_float4*>(fluid->getForces()),
                  (unsigned)fluid->getNumParticles(), reinterpret_cast<fg_float2*>(fluid->getDensityField()),
                  fluid->getHashStart(), fluid->getHashEnd());
}

// Coupled step with the body work on a second stream (one fluid, no stage
// timing).  Dependencies, in the order of HeterogeneousSimulator.cpp:54-127:
//   bodies:  world transform, body grid, bounds        (needs last step's body state)
//   fluid:   hash, sort, cells                         -> [cells]
//   fluid:   density
//   bodies:  forces from the sorted fluid              (after [cells])
//   fluid:   coupled forces + integration              (after the body grid)
//   bodies:  FP64 sums; integration                    (after the fluid forces,
//            which read the body velocities)
// and the fluid stream waits for the body stream at the end, so that one
// step, and a graph captured from it, is complete when the fluid stream is.
void HeterogeneousSimulator::enqueueCoupledOverlapped(Fluid::GPUSPHFluid* f) {
    RigidBody::RigidBodySimulator* rb = RigidBody::RigidBodySimulator::instance();
    enum { kFork, kBodyGrid, kCells, kForces, kJoin };
    rb->startInteraction();
    f->beginUpdate();
    emit(f);
    hip_check(hipEventRecord(mSync[kFork], mStream), "event");
    hip_check(hipStreamWaitEvent(mBodyStream, mSync[kFork], 0), "wait");
    fg_set_stream(mBodyStream);
    rb->setStream(mBodyStream);
    rb->generateCollisionData();
    hip_check(hipEventRecord(mSync[kBodyGrid], mBodyStream), "event");
    fg_set_stream(mStream);
    findFluidNeighbourhood(f);
    hip_check(hipEventRecord(mSync[kCells], mStream), "event");
    calculateFluidDensity(f);
    hip_check(hipStreamWaitEvent(mBodyStream, mSync[kCells], 0), "wait");
    fg_set_stream(mBodyStream);
    rb->calculateRigidBodyForcesFromFluid(f->getOrderedPosition(), f->getOrderedVelocities(), f->getHashStart(),
                                          f->getHashEnd(), (unsigned)f->getNumParticles());
    rb->applyForces();
    fg_set_stream(mStream);
    hip_check(hipStreamWaitEvent(mStream, mSync[kBodyGrid], 0), "wait");
    calculateFluidForces(f, hasStaticBodies(), true);
    if (!mOptions.fused) integrateFluidForces(f);
    hip_check(hipEventRecord(mSync[kForces], mStream), "event");
    hip_check(hipStreamWaitEvent(mBodyStream, mSync[kForces], 0), "wait");
    fg_set_stream(mBodyStream);
    rb->integrate();
    hip_check(hipEventRecord(mSync[kJoin], mBodyStream), "event");
    fg_set_stream(mStream);
    rb->setStream(mStream);
    hip_check(hipStreamWaitEvent(mStream, mSync[kJoin], 0), "wait");
    f->endUpdate();
    rb->endInteraction();
}

// Events: 0 start, 1 hash, 2 sort, 3 cells, 4 density, 5 bodies (collision
// data + forces from the fluid), 6 fluid forces, 7 fluid integrate, 8 body
// sums and integration.
void HeterogeneousSimulator::enqueueStep() {
    const bool t = mOptions.stage_timing;
    RigidBody::RigidBodySimulator* rb = RigidBody::RigidBodySimulator::instance();
    const bool withBodies = coupled();
    const bool withStatic = hasStaticBodies();
    if (withBodies && mFluids.size() != 1)
        throw std::runtime_error("HeterogeneousSimulator: rigid-body coupling supports exactly one fluid");
    if (withBodies && !t && mOptions.overlap_bodies) {
        enqueueCoupledOverlapped(mFluids[0]);
        return;
    }
    if (withBodies) rb->startInteraction();
    for (Fluid::GPUSPHFluid* f : mFluids) {
        f->beginUpdate();
        emit(f);
        if (t) (void)hipEventRecord(mEv[0], mStream);
        findFluidNeighbourhood(f);
        if (t) (void)hipEventRecord(mEv[3], mStream);
        calculateFluidDensity(f);
        if (t) (void)hipEventRecord(mEv[4], mStream);
        if (withBodies) {
            rb->generateCollisionData();
            rb->calculateRigidBodyForcesFromFluid(f->getOrderedPosition(), f->getOrderedVelocities(),
                                                  f->getHashStart(), f->getHashEnd(), (unsigned)f->getNumParticles());
        }
        if (t) (void)hipEventRecord(mEv[5], mStream);
        calculateFluidForces(f, withStatic, withBodies);
        if (t) (void)hipEventRecord(mEv[6], mStream);
        if (!mOptions.fused) integrateFluidForces(f);
        if (t) (void)hipEventRecord(mEv[7], mStream);
        f->endUpdate();
    }
    if (withBodies) {
        rb->applyForces();
        rb->integrate();
        rb->endInteraction();
    }
    if (t) (void)hipEventRecord(mEv[8], mStream);
}

// HeterogeneousSimulator.cpp:54-127
void HeterogeneousSimulator::doTimeStep() {
    StreamScope scope(mStream);
    PrecisionScope precision(mOptions.fused ? mOptions.precision : FG_PRECISION_FP32);
    RigidBody::RigidBodySimulator::instance()->setStream(mStream);
    enqueueStep();
    fg_check("doTimeStep");
    if (mOptions.stage_timing && !mFluids.empty()) {
        hip_check(hipEventSynchronize(mEv[8]), "stage timing");
        float body_a = 0, body_b = 0;
        float* out[4] = {&mTimes.hash_ms, &mTimes.sort_ms, &mTimes.cells_ms, &mTimes.density_ms};
        for (int k = 0; k < 4; k++) (void)hipEventElapsedTime(out[k], mEv[k], mEv[k + 1]);
        (void)hipEventElapsedTime(&body_a, mEv[4], mEv[5]);
        (void)hipEventElapsedTime(&mTimes.forces_ms, mEv[5], mEv[6]);
        (void)hipEventElapsedTime(&mTimes.integrate_ms, mEv[6], mEv[7]);
        (void)hipEventElapsedTime(&body_b, mEv[7], mEv[8]);
        mTimes.bodies_ms = body_a + body_b;
    }
}

void HeterogeneousSimulator::doTimeSteps(int steps) {
    if (steps <= 0) return;
    if (!mOptions.use_graph || mOptions.stage_timing) {
        for (int s = 0; s < steps; s++) doTimeStep();
        return;
    }
    StreamScope scope(mStream);
    PrecisionScope precision(mOptions.fused ? mOptions.precision : FG_PRECISION_FP32);
    RigidBody::RigidBodySimulator::instance()->setStream(mStream);
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

void HeterogeneousSimulator::applyTuning(const StepTuning& t) {
    if (fg_set_tuning(t.sphBlock, t.sortDigitBits) != FG_OK) fg_check("applyTuning");
    fg_set_schedule(t.flatDensity ? 1 : 0, t.flatForces ? 1 : 0);
    fg_set_tiled(t.tiled ? 1 : 0);
}

StepTuning HeterogeneousSimulator::currentTuning() {
    StepTuning t;
    int blk = 256, bits = 8, fd = 0, ff = 0, tl = 0;
    fg_get_tuning(&blk, &bits);
    fg_get_schedule(&fd, &ff);
    fg_get_tiled(&tl);
    t.sphBlock = blk;
    t.sortDigitBits = bits;
    t.flatDensity = fd != 0;
    t.flatForces = ff != 0;
    t.tiled = tl != 0;
    return t;
}

StepTuning HeterogeneousSimulator::autotune(int trialSteps, std::vector<std::pair<StepTuning, float>>* trials) {
    if (trialSteps < 1) trialSteps = 1;
    if (mFluids.empty()) throw std::runtime_error("autotune: no fluid");
    hip_check(hipStreamSynchronize(mStream), "sync");
    // What a step reads from the one before (as saveState keeps it), saved
    // on the device (bodies and inflow counters through the host).
    struct FluidSnap {
        float4 *pos = nullptr, *vel = nullptr;
        float2* rho = nullptr;
        unsigned emit = 0;
    };
    std::vector<FluidSnap> snaps(mFluids.size());
    for (size_t k = 0; k < mFluids.size(); k++) {
        Fluid::GPUSPHFluid* f = mFluids[k];
        const size_t n = (size_t)(f->getNumParticles() > 0 ? f->getNumParticles() : 1);
        FluidSnap& sn = snaps[k];
        hip_check(hipMalloc((void**)&sn.pos, sizeof(float4) * n), "autotune hipMalloc");
        hip_check(hipMalloc((void**)&sn.vel, sizeof(float4) * n), "autotune hipMalloc");
        hip_check(hipMalloc((void**)&sn.rho, sizeof(float2) * n), "autotune hipMalloc");
        hip_check(hipMemcpy(sn.pos, f->getPosition(), sizeof(float4) * n, hipMemcpyDeviceToDevice), "save");
        hip_check(hipMemcpy(sn.vel, f->getVelocity(), sizeof(float4) * n, hipMemcpyDeviceToDevice), "save");
        hip_check(hipMemcpy(sn.rho, f->getDensityField(), sizeof(float2) * n, hipMemcpyDeviceToDevice), "save");
        if (f->hasInflow()) sn.emit = f->readEmitStep();
    }
    RigidBody::RigidBodySimulator* rb = RigidBody::RigidBodySimulator::instance();
    const size_t B = rb->getNumDynamicBodies();
    std::vector<float4> bpos(B), bvel(B), bori(B), bL(B);
    if (B) rb->downloadBodies(bpos.data(), bvel.data(), bori.data(), bL.data());
    auto restore = [&]() {
        hip_check(hipStreamSynchronize(mStream), "sync");
        for (size_t k = 0; k < mFluids.size(); k++) {
            Fluid::GPUSPHFluid* f = mFluids[k];
            const size_t n = (size_t)(f->getNumParticles() > 0 ? f->getNumParticles() : 1);
            const FluidSnap& sn = snaps[k];
            hip_check(hipMemcpy(f->getPosition(), sn.pos, sizeof(float4) * n, hipMemcpyDeviceToDevice), "restore");
            hip_check(hipMemcpy(f->getVelocity(), sn.vel, sizeof(float4) * n, hipMemcpyDeviceToDevice), "restore");
            hip_check(hipMemcpy(f->getDensityField(), sn.rho, sizeof(float2) * n, hipMemcpyDeviceToDevice),
                      "restore");
            if (f->hasInflow()) f->writeEmitStep(sn.emit);
        }
        if (B) rb->uploadBodies(bpos.data(), bvel.data(), bori.data(), bL.data());
        hip_check(hipDeviceSynchronize(), "restore");
    };
    const StepOptions saved = mOptions;
    StepOptions o = mOptions;
    o.use_graph = true;
    o.stage_timing = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    auto time_shape = [&](const StepTuning& t, bool record) -> float {
        applyTuning(t);
        setOptions(o);  // drops the captured graph: the shapes are recorded in it
        restore();
        doTimeSteps(2);  // capture + warm-up
        hip_check(hipEventRecord(e0, mStream), "event");
        doTimeSteps(trialSteps);
        hip_check(hipEventRecord(e1, mStream), "event");
        hip_check(hipEventSynchronize(e1), "event");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
        ms /= (float)trialSteps;
        if (trials && record) trials->push_back({t, ms});
        return ms;
    };
    StepTuning best = currentTuning();
    try {
        (void)time_shape(best, false);  // the clocks of an idle GPU ramp up over the first trial
        // Coordinate descent from the current shape; a candidate replaces the
        // incumbent only when it is faster by more than kTuneMargin, about the
        // spread of repeated trials of one shape.
        constexpr float kTuneMargin = 0.01f;
        float best_ms = time_shape(best, true);
        auto consider = [&](const StepTuning& t) {
            const float ms = time_shape(t, true);
            if (ms < best_ms * (1.0f - kTuneMargin)) {
                best_ms = ms;
                best = t;
            }
        };
        for (int blk : {64, 128, 256, 512}) {
            StepTuning t = best;
            t.sphBlock = blk;
            if (t.sphBlock != best.sphBlock) consider(t);
        }
        const StepTuning base = best;
        for (int sched = 1; sched < 4; sched++) {
            StepTuning t = base;
            t.flatDensity = (sched & 1) != 0;
            t.flatForces = (sched & 2) != 0;
            consider(t);
        }
        {
            StepTuning t = best;
            t.tiled = !best.tiled;
            consider(t);
        }
        for (int bits : {7, 6}) {
            StepTuning t = best;
            t.sortDigitBits = bits;
            consider(t);
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        for (FluidSnap& sn : snaps) {
            (void)hipFree(sn.pos);
            (void)hipFree(sn.vel);
            (void)hipFree(sn.rho);
        }
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    applyTuning(best);
    setOptions(saved);
    restore();
    for (FluidSnap& sn : snaps) {
        (void)hipFree(sn.pos);
        (void)hipFree(sn.vel);
        (void)hipFree(sn.rho);
    }
    return best;
}

}  // namespace Simulators
}  // namespace Simulation
