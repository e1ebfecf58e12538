// fg_sim_capi.cpp -- C-ABI over the C++ HeterogeneousSimulator mirror, for
// callers without a C++ toolchain (bench and tests drive the product path
// through it).  One process-wide simulator, as in the reference.
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fluidgrid/HeterogeneousSimulator.h"
#include "fluidgrid_amd_sim.h"

using namespace Simulation;

namespace {
struct SimState {
    std::vector<std::unique_ptr<Emitter::ArrayEmitter>> emitters;
    std::vector<std::unique_ptr<Fluid::GPUSPHFluid>> fluids;
    std::vector<std::unique_ptr<RigidBody::RigidBody>> bodies;
    void* diagWs = nullptr;     // fg_sim_conservation workspace
    double* diagOut = nullptr;  // [8]
    ~SimState() {
        if (diagWs) (void)hipFree(diagWs);
        if (diagOut) (void)hipFree(diagOut);
    }
};

SimState*& sim() {
    static SimState* s = nullptr;
    return s;
}

int fail(const std::exception& e) {
    fg_record_error(FG_ERR_HIP, e.what());
    return FG_ERR_HIP;
}

Fluid::GPUSPHFluid* fluid_at(int idx) {
    if (!sim() || idx < 0 || idx >= (int)sim()->fluids.size()) return nullptr;
    return sim()->fluids[idx].get();
}
}  // namespace

extern "C" {

int fg_sim_shutdown(void) {
    try {
        Simulators::HeterogeneousSimulator::destroy();
        RigidBody::RigidBodySimulator::destroy();
        delete sim();
        sim() = nullptr;
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_init(const fg_common_fluid_params* common, int key_mode, int math_mode) {
    if (!common) return FG_ERR_INVALID_ARGUMENT;
    try {
        fg_sim_shutdown();
        int rc = fg_set_key_mode(key_mode);
        if (rc) return rc;
        rc = fg_set_math_mode(math_mode);
        if (rc) return rc;
        Application::ApplicationData::instance()->setCommonFluidParams(
            Application::abi_cast<Application::CommonFluidParams>(*common));
        sim() = new SimState();
        (void)Simulators::HeterogeneousSimulator::instance();
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_add_fluid(const fg_fluid_params* params, const fg_float4* positions, const fg_float4* velocities) {
    if (!sim() || !params || (params->max_particles > 0 && !positions)) return -FG_ERR_INVALID_ARGUMENT;
    try {
        auto em = std::make_unique<Emitter::ArrayEmitter>(reinterpret_cast<const float4*>(positions),
                                                          reinterpret_cast<const float4*>(velocities),
                                                          params->max_particles);
        auto fl = std::make_unique<Fluid::GPUSPHFluid>(Application::abi_cast<Application::FluidParams>(*params),
                                                       em.get());
        Simulators::HeterogeneousSimulator::instance()->addSphFluid(fl.get());
        sim()->emitters.push_back(std::move(em));
        sim()->fluids.push_back(std::move(fl));
        return (int)sim()->fluids.size() - 1;
    } catch (const std::exception& e) {
        return -fail(e);
    }
}

int fg_sim_set_options(int fused, int use_graph, int stage_timing) {
    if (!sim()) return FG_ERR_NOT_INITIALIZED;
    Simulators::StepOptions o = Simulators::HeterogeneousSimulator::instance()->options();
    o.fused = fused != 0;
    o.use_graph = use_graph != 0;
    o.stage_timing = stage_timing != 0;
    Simulators::HeterogeneousSimulator::instance()->setOptions(o);
    return FG_OK;
}

int fg_sim_set_overlap(int overlap_bodies) {
    if (!sim()) return FG_ERR_NOT_INITIALIZED;
    Simulators::StepOptions o = Simulators::HeterogeneousSimulator::instance()->options();
    o.overlap_bodies = overlap_bodies != 0;
    Simulators::HeterogeneousSimulator::instance()->setOptions(o);
    return FG_OK;
}

int fg_sim_set_precision(int mode) {
    if (!sim()) return FG_ERR_NOT_INITIALIZED;
    if (mode != FG_PRECISION_FP32 && mode != FG_PRECISION_MIXED) {
        fg_record_error(FG_ERR_INVALID_ARGUMENT, "fg_sim_set_precision: unknown mode");
        return FG_ERR_INVALID_ARGUMENT;
    }
    Simulators::StepOptions o = Simulators::HeterogeneousSimulator::instance()->options();
    o.precision = mode;
    Simulators::HeterogeneousSimulator::instance()->setOptions(o);
    return FG_OK;
}

int fg_sim_autotune(int trialSteps, int* tuning, float* trials, int maxTrials, int* numTrials) {
    if (!sim()) return FG_ERR_NOT_INITIALIZED;
    try {
        std::vector<std::pair<Simulators::StepTuning, float>> log;
        const Simulators::StepTuning t = Simulators::HeterogeneousSimulator::instance()->autotune(trialSteps, &log);
        if (tuning) {
            tuning[0] = t.sphBlock;
            tuning[1] = t.sortDigitBits;
            tuning[2] = t.flatDensity ? 1 : 0;
            tuning[3] = t.flatForces ? 1 : 0;
            tuning[4] = t.tiled ? 1 : 0;
        }
        int k = 0;
        for (; trials && k < maxTrials && k < (int)log.size(); k++) {
            float* r = trials + 6 * k;
            r[0] = (float)log[k].first.sphBlock;
            r[1] = (float)log[k].first.sortDigitBits;
            r[2] = log[k].first.flatDensity ? 1.0f : 0.0f;
            r[3] = log[k].first.flatForces ? 1.0f : 0.0f;
            r[4] = log[k].first.tiled ? 1.0f : 0.0f;
            r[5] = log[k].second;
        }
        if (numTrials) *numTrials = k;
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_step(int steps) {
    if (!sim()) return FG_ERR_NOT_INITIALIZED;
    try {
        Simulators::HeterogeneousSimulator::instance()->doTimeSteps(steps);
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_save_state(const char* path) {
    if (!sim()) return FG_ERR_NOT_INITIALIZED;
    if (!path) return FG_ERR_INVALID_ARGUMENT;
    try {
        Simulators::HeterogeneousSimulator::instance()->saveState(path);
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_load_state(const char* path) {
    if (!sim()) return FG_ERR_NOT_INITIALIZED;
    if (!path) return FG_ERR_INVALID_ARGUMENT;
    try {
        Simulators::HeterogeneousSimulator::instance()->loadState(path);
        return FG_OK;
    } catch (const std::runtime_error& e) {
        // a file that does not fit the scene, or a damaged one
        const std::string msg = e.what();
        if (msg.rfind("loadState: ", 0) == 0) {
            fg_record_error(FG_ERR_INVALID_ARGUMENT, msg.c_str());
            return FG_ERR_INVALID_ARGUMENT;
        }
        return fail(e);
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_write_ply(int idx, const char* path, unsigned int* count) {
    Fluid::GPUSPHFluid* f = fluid_at(idx);
    if (!f || !path) return FG_ERR_INVALID_ARGUMENT;
    try {
        const int n = f->writePly(path);
        if (count) *count = (unsigned)n;
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_synchronize(void) {
    if (!sim()) return FG_ERR_NOT_INITIALIZED;
    hipError_t e = hipStreamSynchronize(Simulators::HeterogeneousSimulator::instance()->stream());
    if (e != hipSuccess) {
        fg_record_error(FG_ERR_HIP, hipGetErrorString(e));
        return FG_ERR_HIP;
    }
    return FG_OK;
}

void* fg_sim_stream(void) {
    if (!sim()) return nullptr;
    return reinterpret_cast<void*>(Simulators::HeterogeneousSimulator::instance()->stream());
}

int fg_sim_download(int idx, fg_float4* positions, fg_float4* velocities) {
    Fluid::GPUSPHFluid* f = fluid_at(idx);
    if (!f) return FG_ERR_INVALID_ARGUMENT;
    try {
        f->downloadState(reinterpret_cast<float4*>(positions), reinterpret_cast<float4*>(velocities));
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_set_inflow(int idx, const fg_inflow_desc* desc) {
    Fluid::GPUSPHFluid* f = fluid_at(idx);
    if (!f || !desc) return FG_ERR_INVALID_ARGUMENT;
    try {
        (void)hipStreamSynchronize(Simulators::HeterogeneousSimulator::instance()->stream());
        f->setInflow(*desc);
        Simulators::HeterogeneousSimulator::instance()->invalidateGraph();
        return FG_OK;
    } catch (const std::invalid_argument& e) {
        fg_record_error(FG_ERR_INVALID_ARGUMENT, e.what());
        return FG_ERR_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_conservation(int idx, double* out8) {
    Fluid::GPUSPHFluid* f = fluid_at(idx);
    if (!f || !out8) return FG_ERR_INVALID_ARGUMENT;
    try {
        SimState* st = sim();
        const size_t bytes = fg_conservation_workspace_size();
        if (!st->diagWs && hipMalloc(&st->diagWs, bytes) != hipSuccess)
            throw std::runtime_error("fg_sim_conservation: hipMalloc failed");
        if (!st->diagOut && hipMalloc(reinterpret_cast<void**>(&st->diagOut), 8 * sizeof(double)) != hipSuccess)
            throw std::runtime_error("fg_sim_conservation: hipMalloc failed");
        hipStream_t stream = Simulators::HeterogeneousSimulator::instance()->stream();
        void* prev = fg_get_stream();
        fg_set_stream(stream);
        const int rc = fg_conservation_f64(reinterpret_cast<fg_float4*>(f->getPosition()),
                                           reinterpret_cast<fg_float4*>(f->getVelocity()),
                                           reinterpret_cast<fg_float4*>(f->getOrderedPosition()),
                                           reinterpret_cast<fg_float2*>(f->getDensityField()),
                                           (unsigned)f->getNumParticles(), f->getMass(), f->getRestDensity(),
                                           f->getGasConstant(), st->diagOut, st->diagWs, bytes);
        fg_set_stream(prev);
        if (rc != FG_OK) return rc;
        if (hipStreamSynchronize(stream) != hipSuccess ||
            hipMemcpy(out8, st->diagOut, 8 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("fg_sim_conservation: copy failed");
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_upload(int idx, const fg_float4* positions, const fg_float4* velocities) {
    Fluid::GPUSPHFluid* f = fluid_at(idx);
    if (!f) return FG_ERR_INVALID_ARGUMENT;
    try {
        (void)hipStreamSynchronize(Simulators::HeterogeneousSimulator::instance()->stream());
        f->uploadState(reinterpret_cast<const float4*>(positions), reinterpret_cast<const float4*>(velocities));
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_fluid_buffers(int idx, fg_fluid_buffers* out) {
    Fluid::GPUSPHFluid* f = fluid_at(idx);
    if (!f || !out) return FG_ERR_INVALID_ARGUMENT;
    out->hashKey = f->getHashKey();
    out->particleId = f->getParticleID();
    out->gridStart = f->getHashStart();
    out->gridEnd = f->getHashEnd();
    out->position = f->getPosition();
    out->orderedPosition = f->getOrderedPosition();
    out->velocity = f->getVelocity();
    out->orderedVelocity = f->getOrderedVelocities();
    out->densityPressure = f->getDensityField();
    out->forces = f->getForces();
    out->numParticles = (unsigned)f->getNumParticles();
    out->numBuckets = f->getNumBuckets();
    return FG_OK;
}

int fg_sim_add_bodies(const fg_float4* worldSpheres, const unsigned int* bodyOffsets, unsigned int numBodies,
                      const double* masses, float sphereRadius, const fg_common_rigid_body_params* rigidParams,
                      float springForce, float dampingForce) {
    if (!sim() || !worldSpheres || !bodyOffsets || !masses || !rigidParams) return FG_ERR_INVALID_ARGUMENT;
    try {
        Application::ApplicationData* ad = Application::ApplicationData::instance();
        ad->setCommonRigidBodyParams(Application::abi_cast<Application::CommonRigidBodyParams>(*rigidParams));
        Application::SimulationParams sp{dampingForce, springForce};
        ad->setSimulationParams(sp);
        RigidBody::RigidBodySimulator* rb = RigidBody::RigidBodySimulator::instance();
        for (unsigned b = 0; b < numBodies; b++) {
            std::vector<float4> s;
            for (unsigned i = bodyOffsets[b]; i < bodyOffsets[b + 1]; i++)
                s.push_back(make_float4(worldSpheres[i].x, worldSpheres[i].y, worldSpheres[i].z, 0.0f));
            sim()->bodies.push_back(
                std::make_unique<RigidBody::RigidBody>(false, (float)masses[b], sphereRadius, std::move(s)));
            rb->addElement(sim()->bodies.back().get());
        }
        rb->setStream(Simulators::HeterogeneousSimulator::instance()->stream());
        rb->reset();
        Simulators::HeterogeneousSimulator::instance()->invalidateGraph();
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_add_static_spheres(const fg_float4* worldSpheres, unsigned int numSpheres, float sphereRadius,
                              const fg_common_rigid_body_params* rigidParams, float springForce,
                              float dampingForce) {
    if (!sim() || !worldSpheres || !rigidParams || numSpheres == 0) return FG_ERR_INVALID_ARGUMENT;
    try {
        Application::ApplicationData* ad = Application::ApplicationData::instance();
        ad->setCommonRigidBodyParams(Application::abi_cast<Application::CommonRigidBodyParams>(*rigidParams));
        Application::SimulationParams sp{dampingForce, springForce};
        ad->setSimulationParams(sp);
        std::vector<float4> s;
        for (unsigned i = 0; i < numSpheres; i++)
            s.push_back(make_float4(worldSpheres[i].x, worldSpheres[i].y, worldSpheres[i].z, 0.0f));
        sim()->bodies.push_back(std::make_unique<RigidBody::RigidBody>(true, 0.0f, sphereRadius, std::move(s)));
        RigidBody::RigidBodySimulator* rb = RigidBody::RigidBodySimulator::instance();
        rb->addElement(sim()->bodies.back().get());
        rb->setStream(Simulators::HeterogeneousSimulator::instance()->stream());
        rb->reset();
        Simulators::HeterogeneousSimulator::instance()->invalidateGraph();
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_static_grid(fg_body_grid* out) {
    if (!sim()) return FG_ERR_NOT_INITIALIZED;
    if (!out) return FG_ERR_INVALID_ARGUMENT;
    try {
        const Application::CommonFluidParams& cp = Application::ApplicationData::instance()->getCommonFluidParams();
        *out = RigidBody::RigidBodySimulator::instance()->staticGrid(2.0f * cp.pointRadius);
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_body_state(fg_float4* positions, fg_float4* velocities, fg_float4* orientations,
                      fg_float4* angularMomenta) {
    if (!sim()) return FG_ERR_NOT_INITIALIZED;
    try {
        RigidBody::RigidBodySimulator::instance()->downloadBodies(
            reinterpret_cast<float4*>(positions), reinterpret_cast<float4*>(velocities),
            reinterpret_cast<float4*>(orientations), reinterpret_cast<float4*>(angularMomenta));
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_body_sums(double* forces4, double* torques4) {
    if (!sim()) return FG_ERR_NOT_INITIALIZED;
    try {
        RigidBody::RigidBodySimulator::instance()->downloadSums(forces4, torques4);
        return FG_OK;
    } catch (const std::exception& e) {
        return fail(e);
    }
}

int fg_sim_body_mass_properties(double* centreOfMass3, double* inertia9) {
    if (!sim() || !centreOfMass3 || !inertia9) return FG_ERR_INVALID_ARGUMENT;
    RigidBody::RigidBodySimulator::instance()->downloadMassProperties(centreOfMass3, inertia9);
    return FG_OK;
}

int fg_sim_stage_times(float* out6) {
    if (!sim() || !out6) return FG_ERR_INVALID_ARGUMENT;
    const auto& t = Simulators::HeterogeneousSimulator::instance()->lastStageTimes();
    out6[0] = t.hash_ms;
    out6[1] = t.sort_ms;
    out6[2] = t.cells_ms;
    out6[3] = t.density_ms;
    out6[4] = t.forces_ms;
    out6[5] = t.integrate_ms;
    out6[6] = t.bodies_ms;
    return FG_OK;
}

}  // extern "C"
