// fg_batch_capi.cpp -- C-ABI over BatchedSimulator (one process-wide batch).
#include <memory>
#include <stdexcept>
#include <vector>

#include "fluidgrid/BatchedSimulator.h"
#include "fluidgrid_amd_batch.h"

using namespace Simulation;

namespace {
std::unique_ptr<Simulators::BatchedSimulator>& batch() {
    static std::unique_ptr<Simulators::BatchedSimulator> b;
    return b;
}

int fail(const std::exception& e) {
    int code = FG_ERR_HIP;
    if (dynamic_cast<const Simulators::UnsupportedScene*>(&e))
        code = FG_ERR_UNSUPPORTED;
    else if (dynamic_cast<const std::invalid_argument*>(&e))
        code = FG_ERR_INVALID_ARGUMENT;
    fg_record_error(code, e.what());
    return code;
}
}  // namespace

extern "C" {

int fg_bsim_create(unsigned int numScenes, const fg_common_fluid_params* commons, const fg_fluid_params* fluids,
                   const unsigned int* counts, const fg_float4* positions, const fg_float4* velocities) {
    if (numScenes == 0 || !commons || !fluids || !counts || !positions || !velocities) {
        fg_record_error(FG_ERR_INVALID_ARGUMENT, "fg_bsim_create: null pointer or no scenes");
        return FG_ERR_INVALID_ARGUMENT;
    }
    try {
        batch().reset();
        std::vector<Simulators::BatchScene> scenes(numScenes);
        size_t off = 0;
        for (unsigned s = 0; s < numScenes; s++) {
            scenes[s].common = Application::abi_cast<Application::CommonFluidParams>(commons[s]);
            scenes[s].fluid = Application::abi_cast<Application::FluidParams>(fluids[s]);
            const float4* p = reinterpret_cast<const float4*>(positions) + off;
            const float4* v = reinterpret_cast<const float4*>(velocities) + off;
            scenes[s].positions.assign(p, p + counts[s]);
            scenes[s].velocities.assign(v, v + counts[s]);
            off += counts[s];
        }
        batch() = std::make_unique<Simulators::BatchedSimulator>(scenes);
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_bsim_add_bodies(unsigned int scene, const fg_float4* worldSpheres, const unsigned int* bodyOffsets,
                       unsigned int numBodies, const double* masses, float sphereRadius,
                       const fg_common_rigid_body_params* rigidParams, float springForce, float dampingForce) {
    if (!batch()) return FG_ERR_NOT_INITIALIZED;
    if (!worldSpheres || !bodyOffsets || !masses || !rigidParams || numBodies == 0) {
        fg_record_error(FG_ERR_INVALID_ARGUMENT, "fg_bsim_add_bodies: null pointer or no bodies");
        return FG_ERR_INVALID_ARGUMENT;
    }
    try {
        Simulators::BatchBodies b;
        b.offsets.assign(bodyOffsets, bodyOffsets + numBodies + 1);
        const float4* w = reinterpret_cast<const float4*>(worldSpheres);
        b.worldSpheres.assign(w, w + bodyOffsets[numBodies]);
        b.masses.assign(masses, masses + numBodies);
        b.sphereRadius = sphereRadius;
        b.rigid = Application::abi_cast<Application::CommonRigidBodyParams>(*rigidParams);
        b.springForce = springForce;
        b.dampingForce = dampingForce;
        batch()->addBodies(scene, b);
        return FG_OK;
    } catch (const std::invalid_argument& e) {
        fg_record_error(FG_ERR_INVALID_ARGUMENT, e.what());
        return FG_ERR_INVALID_ARGUMENT;
    } catch (const std::logic_error& e) {
        fg_record_error(FG_ERR_INVALID_ARGUMENT, e.what());
        return FG_ERR_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_bsim_body_state(unsigned int* numBodies, fg_float4* positions, fg_float4* velocities, fg_float4* orientations,
                       fg_float4* angularMomenta) {
    if (!batch()) return FG_ERR_NOT_INITIALIZED;
    try {
        if (numBodies) *numBodies = batch()->numBodies();
        batch()->downloadBodies(reinterpret_cast<float4*>(positions), reinterpret_cast<float4*>(velocities),
                                reinterpret_cast<float4*>(orientations), reinterpret_cast<float4*>(angularMomenta));
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_bsim_set_options(int use_graph, int stage_timing) {
    if (!batch()) return FG_ERR_NOT_INITIALIZED;
    batch()->setUseGraph(use_graph != 0);
    batch()->setStageTiming(stage_timing != 0);
    return FG_OK;
}

int fg_bsim_set_overlap(int overlap_bodies) {
    if (!batch()) return FG_ERR_NOT_INITIALIZED;
    batch()->setOverlap(overlap_bodies != 0);
    return FG_OK;
}

int fg_bsim_step(int steps) {
    if (!batch()) return FG_ERR_NOT_INITIALIZED;
    try {
        batch()->doTimeSteps(steps);
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

void* fg_bsim_stream(void) { return batch() ? reinterpret_cast<void*>(batch()->stream()) : nullptr; }

int fg_bsim_synchronize(void) {
    if (!batch()) return FG_ERR_NOT_INITIALIZED;
    try {
        batch()->synchronize();
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_bsim_download(fg_float4* positions, fg_float4* velocities) {
    if (!batch()) return FG_ERR_NOT_INITIALIZED;
    try {
        batch()->download(reinterpret_cast<float4*>(positions), reinterpret_cast<float4*>(velocities));
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_bsim_download_density(fg_float2* rho_p, unsigned int* particleId) {
    if (!batch()) return FG_ERR_NOT_INITIALIZED;
    try {
        batch()->downloadDensity(reinterpret_cast<float2*>(rho_p), particleId);
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_bsim_stage_times(float* out5) {
    if (!batch() || !out5) return FG_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < 5; k++) out5[k] = batch()->stageTimes()[k];
    return FG_OK;
}

int fg_bsim_info(unsigned int* numParticles, unsigned int* numBuckets, int* keyBits) {
    if (!batch()) return FG_ERR_NOT_INITIALIZED;
    if (numParticles) *numParticles = batch()->numParticles();
    if (numBuckets) *numBuckets = batch()->numBuckets();
    if (keyBits) *keyBits = batch()->keyBits();
    return FG_OK;
}

int fg_bsim_autotune(int trialSteps, int* tuning, float* trials, int maxTrials, int* numTrials) {
    if (!batch()) return FG_ERR_NOT_INITIALIZED;
    try {
        std::vector<std::pair<Simulators::BatchTuning, float>> log;
        const Simulators::BatchTuning t = batch()->autotune(trialSteps, &log);
        if (tuning) {
            tuning[0] = t.sphBlock;
            tuning[1] = t.sortDigitBits;
            tuning[2] = t.flatDensity ? 1 : 0;
            tuning[3] = t.flatForces ? 1 : 0;
            tuning[4] = t.queueItems;
            tuning[5] = t.deferContact ? 1 : 0;
        }
        int k = 0;
        for (; trials && k < maxTrials && k < (int)log.size(); k++) {
            float* r = trials + 7 * k;
            r[0] = (float)log[k].first.sphBlock;
            r[1] = (float)log[k].first.sortDigitBits;
            r[2] = log[k].first.flatDensity ? 1.0f : 0.0f;
            r[3] = log[k].first.flatForces ? 1.0f : 0.0f;
            r[4] = (float)log[k].first.queueItems;
            r[5] = log[k].first.deferContact ? 1.0f : 0.0f;
            r[6] = log[k].second;
        }
        if (numTrials) *numTrials = k;
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_bsim_set_tuning(int sphBlock, int sortDigitBits, int flatDensity, int flatForces, int queueItems,
                       int deferContact) {
    if (!batch()) return FG_ERR_NOT_INITIALIZED;
    try {
        Simulators::BatchTuning t;
        t.sphBlock = sphBlock;
        t.sortDigitBits = sortDigitBits;
        t.flatDensity = flatDensity != 0;
        t.flatForces = flatForces != 0;
        t.queueItems = queueItems;
        t.deferContact = deferContact != 0;
        batch()->setTuning(t);
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_bsim_destroy(void) {
    batch().reset();
    return FG_OK;
}

}  // extern "C"
