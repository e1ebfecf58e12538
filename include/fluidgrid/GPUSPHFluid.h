// GPUSPHFluid.h -- owner of one fluid's device buffers, with the getters of
// the reference's GPUSPHFluid.h:117-132 so HeterogeneousSimulator can wire
// the launchers the same way.  No GL: positions live in a plain device array
// instead of a mapped VBO (GPUSPHFluid.cpp:17-41, GLCudaBuffer.h:32-78).
#pragma once

#include <string>

#include <vector>

#include "ApplicationData.h"

namespace Simulation {
namespace Fluid {
class GPUSPHFluid;
}

namespace Emitter {

// Emitter.h:26-33
class IEmitter {
public:
    virtual ~IEmitter() = default;
    virtual void update(Simulation::Fluid::GPUSPHFluid* fluid) = 0;
    virtual void reset(float4* positions, int count) = 0;
};

// Emitter.h:36-55 / Emitter.cpp:16-42: y-layers of an x/z lattice filled
// with float-accumulated spacing; slots beyond the block are (0,0,0,-1).
class BlockEmitter : public IEmitter {
public:
    BlockEmitter(float3 blockMin, float3 blockMax, float spacing)
        : mMin(blockMin), mMax(blockMax), mSpacing(spacing) {}
    void reset(float4* positions, int count) override;
    void update(Simulation::Fluid::GPUSPHFluid*) override {}

private:
    float3 mMin, mMax;
    float mSpacing;
};

// Initial positions (and optionally velocities) taken from host arrays.
class ArrayEmitter : public IEmitter {
public:
    ArrayEmitter(const float4* positions, const float4* velocities, int count);
    void reset(float4* positions, int count) override;
    void update(Simulation::Fluid::GPUSPHFluid*) override {}
    const float4* velocities() const { return mVel.empty() ? nullptr : mVel.data(); }

private:
    std::vector<float4> mPos, mVel;
};

}  // namespace Emitter

namespace Fluid {

class GPUSPHFluid {
public:
    // GPUSPHFluid.cpp:9-42: allocates the particle/grid buffers, zeroes the
    // velocities and fills positions from the emitter.
    GPUSPHFluid(Application::FluidParams fluidParams, Emitter::IEmitter* emitter);
    ~GPUSPHFluid();
    GPUSPHFluid(const GPUSPHFluid&) = delete;
    GPUSPHFluid& operator=(const GPUSPHFluid&) = delete;

    void beginUpdate() {}
    void endUpdate() {}

    unsigned int* getHashKey() { return d_hashKey; }
    unsigned int* getParticleID() { return d_particleId; }
    unsigned int* getHashStart() { return d_gridPosStart; }
    unsigned int* getHashEnd() { return d_gridPosEnd; }
    float4* getPosition() { return d_position; }
    float4* getOrderedPosition() { return d_ordered_position; }
    float4* getOrderedVelocities() { return d_ordered_vel; }
    float4* getVelocity() { return d_velocity; }
    float4* getForces() { return d_forces; }
    float2* getDensityField() { return d_mass_pressure; }
    int getNumParticles() const { return mFluidParams.max_particles; }
    float getGasConstant() const { return mFluidParams.gasConstant; }
    float getRestDensity() const { return mFluidParams.restDensity; }
    float getMass() const { return mFluidParams.mass; }
    float getViscosity() const { return mFluidParams.viscosity; }
    unsigned int getNumBuckets() const { return mNumBuckets; }

    // GPUSPHFluid.cpp:91-98 copies the first max_particles *bytes* of the
    // ordered positions; this copies all max_particles float4 positions in
    // particle order into a malloc'ed array the caller frees.
    int copyPositionToCPU(float4** particlePositions);

    // Extensions: state transfer and the persistent sort workspace.
    void uploadState(const float4* pos, const float4* vel);
    void downloadState(float4* pos, float4* vel);
    // Inflow emission (IEmitter::update, a no-op in the reference): the
    // particle array's reserve slots are emitted layer by layer at the start
    // of each step (fg_emit_inflow), from a step counter kept on the device.
    void setInflow(const fg_inflow_desc& desc);
    bool hasInflow() const { return d_emitStep != nullptr; }
    void emitInflow(float dt);
    const fg_inflow_desc& getInflow() const { return mInflow; }
    // Checkpoint support: the inflow step counter (0 without an inflow) and
    // the sorted (rho, p) buffer, whose entries at inactive particles carry
    // over from step to step (the density pass leaves them as they are and
    // neighbours read them).  Both synchronise the device.
    unsigned int readEmitStep() const;
    void writeEmitStep(unsigned int step);
    void downloadDensity(float2* rho_p) const;
    void uploadDensity(const float2* rho_p);
    // Point cloud for visualisation (the reference draws the positions
    // from a GL buffer, GPUSPHFluid.cpp:66-89): a binary little-endian PLY
    // file with one vertex per active particle (w > 0), properties x y z
    // (particle order), vx vy vz, and density and pressure of the last step
    // (looked up through the sorted particle ids).  Returns the vertex count.
    int writePly(const std::string& path) const;
    void* getSortWorkspace() { return d_sortWorkspace; }
    unsigned long long* getNeighborMasks() { return d_neighborMasks; }
    size_t getSortWorkspaceBytes() const { return mSortWorkspaceBytes; }

    Application::FluidParams mFluidParams;
    Emitter::IEmitter* mEmitter;

private:
    unsigned int mNumBuckets = 0;
    unsigned int* d_particleId = nullptr;
    unsigned int* d_hashKey = nullptr;
    unsigned int* d_gridPosStart = nullptr;
    unsigned int* d_gridPosEnd = nullptr;
    float4* d_position = nullptr;
    float4* d_ordered_position = nullptr;
    float4* d_forces = nullptr;
    float2* d_mass_pressure = nullptr;
    float4* d_velocity = nullptr;
    float4* d_ordered_vel = nullptr;
    void* d_sortWorkspace = nullptr;
    unsigned long long* d_neighborMasks = nullptr;
    size_t mSortWorkspaceBytes = 0;
    fg_inflow_desc mInflow{};
    unsigned int* d_emitStep = nullptr;
};

}  // namespace Fluid
}  // namespace Simulation
