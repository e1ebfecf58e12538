// BatchedSimulator.cpp -- host side of the batched-scene step (fg_batch.hip).
#include "fluidgrid/BatchedSimulator.h"

#include <cmath>
#include <stdexcept>
#include <string>

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
    hip_check(hipStreamCreateWithFlags(&mStream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : mEv) hip_check(hipEventCreate(&e), "hipEventCreate");
    int blk = 256, bits = 8, fd = 0, ff = 0;
    fg_get_tuning(&blk, &bits);
    fg_get_schedule(&fd, &ff);
    mTuning.sphBlock = blk;
    mTuning.sortDigitBits = bits;
    mTuning.flatDensity = fd != 0;
    mTuning.flatForces = ff != 0;
}

BatchedSimulator::~BatchedSimulator() {
    invalidateGraph();
    if (mStream) (void)hipStreamSynchronize(mStream);
    for (auto& e : mEv)
        if (e) (void)hipEventDestroy(e);
    if (mStream) (void)hipStreamDestroy(mStream);
    dfree(d_descs); dfree(d_sceneOf); dfree(d_pos); dfree(d_vel); dfree(d_spos); dfree(d_svel); dfree(d_rho);
    dfree(d_key); dfree(d_id); dfree(d_start); dfree(d_end); dfree(d_masks);
    if (d_sortWs) (void)hipFree(d_sortWs);
}

void BatchedSimulator::invalidateGraph() {
    if (mGraphExec) (void)hipGraphExecDestroy(mGraphExec);
    if (mGraph) (void)hipGraphDestroy(mGraph);
    mGraphExec = nullptr;
    mGraph = nullptr;
}

void BatchedSimulator::setUseGraph(bool on) { mUseGraph = on; }
void BatchedSimulator::setStageTiming(bool on) { mTiming = on; }

void BatchedSimulator::setTuning(const BatchTuning& t) {
    if (fg_set_tuning(t.sphBlock, t.sortDigitBits) != FG_OK) fg_check("setTuning");
    fg_set_schedule(t.flatDensity ? 1 : 0, t.flatForces ? 1 : 0);
    mTuning = t;
    invalidateGraph();  // the launch shapes are recorded in the graph
}

// Stage order of HeterogeneousSimulator::doTimeStep for the fused fast path.
void BatchedSimulator::enqueueStep() {
    const unsigned S = (unsigned)mDescs.size();
    if (mTiming) (void)hipEventRecord(mEv[0], mStream);
    fg_batch_hash(d_descs, S, d_sceneOf, reinterpret_cast<fg_float4*>(d_pos), d_key, mN);
    if (mTiming) (void)hipEventRecord(mEv[1], mStream);
    fg_sort_pairs_ws(d_key, d_id, mN, mKeyBits, d_sortWs, mSortWsBytes, 1);
    if (mTiming) (void)hipEventRecord(mEv[2], mStream);
    gFindCellOffsetOrderingPosVel(d_start, d_end, d_key, d_id, reinterpret_cast<fg_float4*>(d_pos),
                                  reinterpret_cast<fg_float4*>(d_spos), reinterpret_cast<fg_float4*>(d_vel),
                                  reinterpret_cast<fg_float4*>(d_svel), mN, mBuckets);
    if (mTiming) (void)hipEventRecord(mEv[3], mStream);
    fg_batch_density_nbr(d_descs, S, d_sceneOf, reinterpret_cast<fg_float4*>(d_spos), reinterpret_cast<fg_float2*>(d_rho),
                         d_start, d_end, d_masks, mN);
    if (mTiming) (void)hipEventRecord(mEv[4], mStream);
    fg_batch_forces_integrate_nbr(d_descs, S, d_sceneOf, reinterpret_cast<fg_float4*>(d_spos),
                                  reinterpret_cast<fg_float4*>(d_svel), reinterpret_cast<fg_float2*>(d_rho), d_start,
                                  d_end, d_masks, reinterpret_cast<fg_float4*>(d_vel),
                                  reinterpret_cast<fg_float4*>(d_pos), d_id, mN);
    if (mTiming) (void)hipEventRecord(mEv[5], mStream);
}

void BatchedSimulator::doTimeStep() {
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
    synchronize();
    // Save the state so that the trials leave the simulation where it was.
    float4 *bp = dalloc<float4>(mN), *bv = dalloc<float4>(mN);
    hip_check(hipMemcpy(bp, d_pos, sizeof(float4) * mN, hipMemcpyDeviceToDevice), "save");
    hip_check(hipMemcpy(bv, d_vel, sizeof(float4) * mN, hipMemcpyDeviceToDevice), "save");
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
    // Coordinate descent: workgroup size first, then the digit width.
    BatchTuning best = mTuning;
    float best_ms = 1e30f;
    for (int blk : {64, 128, 256, 512}) {
        BatchTuning t = best;
        t.sphBlock = blk;
        const float ms = time_shape(t);
        if (ms < best_ms) { best_ms = ms; best = t; }
    }
    for (int sched = 0; sched < 4; sched++) {
        BatchTuning t = best;
        t.flatDensity = (sched & 1) != 0;
        t.flatForces = (sched & 2) != 0;
        if (t.flatDensity == best.flatDensity && t.flatForces == best.flatForces) continue;
        const float ms = time_shape(t);
        if (ms < best_ms) { best_ms = ms; best = t; }
    }
    for (int bits : {8, 7, 6}) {
        BatchTuning t = best;
        t.sortDigitBits = bits;
        if (t.sortDigitBits == best.sortDigitBits && !trials) continue;
        const float ms = time_shape(t);
        if (ms < best_ms) { best_ms = ms; best = t; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    setTuning(best);
    hip_check(hipMemcpyAsync(d_pos, bp, sizeof(float4) * mN, hipMemcpyDeviceToDevice, mStream), "restore");
    hip_check(hipMemcpyAsync(d_vel, bv, sizeof(float4) * mN, hipMemcpyDeviceToDevice, mStream), "restore");
    hip_check(hipStreamSynchronize(mStream), "restore");
    dfree(bp);
    dfree(bv);
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
