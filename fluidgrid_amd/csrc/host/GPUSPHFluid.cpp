// GPUSPHFluid.cpp -- device buffers of one fluid (GPUSPHFluid.cpp:9-56).
#include "fluidgrid/GPUSPHFluid.h"

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

namespace Simulation {
namespace Fluid {

namespace {
void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
T* dalloc(size_t count, const char* what) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, sizeof(T) * (count ? count : 1)), what);
    return static_cast<T*>(p);
}

int key_bits_for(unsigned int numBuckets) {
    int b = 1;
    while (b < 32 && ((numBuckets - 1u) >> b) != 0) b++;
    return b;
}
}  // namespace

GPUSPHFluid::GPUSPHFluid(Application::FluidParams fluidParams, Emitter::IEmitter* emitter)
    : mFluidParams(fluidParams), mEmitter(emitter) {
    const Application::CommonFluidParams& cp = Application::ApplicationData::instance()->getCommonFluidParams();
    const int n = mFluidParams.max_particles;
    if (n < 0) throw std::invalid_argument("GPUSPHFluid: max_particles < 0");
    const long long nb = (long long)cp.bucket_size.x * cp.bucket_size.y * cp.bucket_size.z;
    if (nb <= 0 || nb > 0xFFFFFFFFll) throw std::invalid_argument("GPUSPHFluid: invalid bucket_size");
    mNumBuckets = (unsigned int)nb;

    d_particleId = dalloc<unsigned int>(n, "hipMalloc particleId");
    d_hashKey = dalloc<unsigned int>(n, "hipMalloc hashKey");
    d_gridPosStart = dalloc<unsigned int>(mNumBuckets, "hipMalloc gridPosStart");
    d_gridPosEnd = dalloc<unsigned int>(mNumBuckets, "hipMalloc gridPosEnd");
    d_position = dalloc<float4>(n, "hipMalloc position");
    d_ordered_position = dalloc<float4>(n, "hipMalloc ordered_position");
    d_forces = dalloc<float4>(n, "hipMalloc forces");
    d_mass_pressure = dalloc<float2>(n, "hipMalloc mass_pressure");
    d_velocity = dalloc<float4>(n, "hipMalloc velocity");
    d_ordered_vel = dalloc<float4>(n, "hipMalloc ordered_vel");
    mSortWorkspaceBytes = fg_sort_workspace_size((unsigned)n, key_bits_for(mNumBuckets));
    hip_check(hipMalloc(&d_sortWorkspace, mSortWorkspaceBytes ? mSortWorkspaceBytes : 256), "hipMalloc sort ws");
    hip_check(hipMalloc((void**)&d_neighborMasks, fg_neighbor_mask_bytes((unsigned)n) + 256), "hipMalloc masks");

    // GPUSPHFluid.cpp:27-29 zeroes the velocities; the density field is
    // zeroed here as well (the reference leaves it uninitialised).
    hip_check(hipMemset(d_velocity, 0, sizeof(float4) * n), "hipMemset velocity");
    hip_check(hipMemset(d_ordered_vel, 0, sizeof(float4) * n), "hipMemset ordered_vel");
    hip_check(hipMemset(d_mass_pressure, 0, sizeof(float2) * n), "hipMemset mass_pressure");
    hip_check(hipMemset(d_forces, 0, sizeof(float4) * n), "hipMemset forces");
    hip_check(hipMemset(d_gridPosEnd, 0, sizeof(unsigned int) * mNumBuckets), "hipMemset gridPosEnd");
    // Every cell empty and no key occupying one: the grid state the fused
    // step keeps between steps (it clears only the cells the last sorted keys
    // occupied, fg_clear_occupied_cells, instead of filling every bucket).
    hip_check(hipMemset(d_gridPosStart, 0xFF, sizeof(unsigned int) * mNumBuckets), "hipMemset gridPosStart");
    hip_check(hipMemset(d_hashKey, 0xFF, sizeof(unsigned int) * (n > 0 ? n : 1)), "hipMemset hashKey");

    std::vector<float4> host(n > 0 ? n : 1);
    if (mEmitter) mEmitter->reset(host.data(), n);
    if (n > 0) hip_check(hipMemcpy(d_position, host.data(), sizeof(float4) * n, hipMemcpyHostToDevice), "upload");
    if (auto* ae = dynamic_cast<Emitter::ArrayEmitter*>(mEmitter)) {
        if (ae->velocities() && n > 0)
            hip_check(hipMemcpy(d_velocity, ae->velocities(), sizeof(float4) * n, hipMemcpyHostToDevice), "upload");
    }
}

GPUSPHFluid::~GPUSPHFluid() {
    (void)hipFree(d_particleId);
    (void)hipFree(d_hashKey);
    (void)hipFree(d_gridPosStart);
    (void)hipFree(d_gridPosEnd);
    (void)hipFree(d_position);
    (void)hipFree(d_ordered_position);
    (void)hipFree(d_forces);
    (void)hipFree(d_mass_pressure);
    (void)hipFree(d_velocity);
    (void)hipFree(d_ordered_vel);
    (void)hipFree(d_sortWorkspace);
    (void)hipFree(d_neighborMasks);
    (void)hipFree(d_emitStep);
}

int GPUSPHFluid::copyPositionToCPU(float4** particlePositions) {
    const int n = mFluidParams.max_particles;
    *particlePositions = static_cast<float4*>(std::malloc(sizeof(float4) * (n > 0 ? n : 1)));
    if (n > 0) hip_check(hipMemcpy(*particlePositions, d_position, sizeof(float4) * n, hipMemcpyDeviceToHost),
                         "copyPositionToCPU");
    return n;
}

void GPUSPHFluid::uploadState(const float4* pos, const float4* vel) {
    const int n = mFluidParams.max_particles;
    if (n <= 0) return;
    if (pos) hip_check(hipMemcpy(d_position, pos, sizeof(float4) * n, hipMemcpyHostToDevice), "uploadState");
    if (vel) hip_check(hipMemcpy(d_velocity, vel, sizeof(float4) * n, hipMemcpyHostToDevice), "uploadState");
}

void GPUSPHFluid::downloadState(float4* pos, float4* vel) {
    const int n = mFluidParams.max_particles;
    if (n <= 0) return;
    hip_check(hipDeviceSynchronize(), "downloadState sync");
    if (pos) hip_check(hipMemcpy(pos, d_position, sizeof(float4) * n, hipMemcpyDeviceToHost), "downloadState");
    if (vel) hip_check(hipMemcpy(vel, d_velocity, sizeof(float4) * n, hipMemcpyDeviceToHost), "downloadState");
}

}  // namespace Fluid
}  // namespace Simulation

namespace Simulation {
namespace Fluid {

void GPUSPHFluid::setInflow(const fg_inflow_desc& desc) {
    const uint64_t reserve = (uint64_t)desc.nu * desc.nv * desc.numLayers;
    if ((uint64_t)desc.firstSlot + reserve > (uint64_t)getNumParticles())
        throw std::invalid_argument("GPUSPHFluid::setInflow: reserve slots beyond max_particles");
    if (desc.nu == 0 || desc.nv == 0 || !(desc.spacing > 0.0f) || !(desc.speed > 0.0f))
        throw std::invalid_argument("GPUSPHFluid::setInflow: empty window or non-positive spacing/speed");
    // At most one layer is due per step (fg_emit_inflow refuses dt > period).
    const Application::CommonFluidParams& cp = Application::ApplicationData::instance()->getCommonFluidParams();
    if ((double)cp.deltaTimeIntegration > fg_inflow_period(&desc, cp.simulation_scale))
        throw std::invalid_argument("GPUSPHFluid::setInflow: the time step exceeds the layer period "
                                    "spacing * simulation_scale / speed");
    mInflow = desc;
    if (!d_emitStep && hipMalloc(reinterpret_cast<void**>(&d_emitStep), sizeof(unsigned int)) != hipSuccess)
        throw std::runtime_error("GPUSPHFluid::setInflow: hipMalloc failed");
    if (hipMemset(d_emitStep, 0, sizeof(unsigned int)) != hipSuccess)
        throw std::runtime_error("GPUSPHFluid::setInflow: hipMemset failed");
}

unsigned int GPUSPHFluid::readEmitStep() const {
    unsigned int k = 0;
    if (!d_emitStep) return 0;
    hip_check(hipDeviceSynchronize(), "readEmitStep sync");
    hip_check(hipMemcpy(&k, d_emitStep, sizeof(k), hipMemcpyDeviceToHost), "readEmitStep");
    return k;
}

void GPUSPHFluid::writeEmitStep(unsigned int step) {
    if (!d_emitStep) {
        if (step) throw std::invalid_argument("GPUSPHFluid::writeEmitStep: the fluid has no inflow");
        return;
    }
    hip_check(hipDeviceSynchronize(), "writeEmitStep sync");
    hip_check(hipMemcpy(d_emitStep, &step, sizeof(step), hipMemcpyHostToDevice), "writeEmitStep");
}

void GPUSPHFluid::downloadDensity(float2* rho_p) const {
    const int n = mFluidParams.max_particles;
    if (n <= 0 || !rho_p) return;
    hip_check(hipDeviceSynchronize(), "downloadDensity sync");
    hip_check(hipMemcpy(rho_p, d_mass_pressure, sizeof(float2) * n, hipMemcpyDeviceToHost), "downloadDensity");
}

void GPUSPHFluid::uploadDensity(const float2* rho_p) {
    const int n = mFluidParams.max_particles;
    if (n <= 0 || !rho_p) return;
    hip_check(hipDeviceSynchronize(), "uploadDensity sync");
    hip_check(hipMemcpy(d_mass_pressure, rho_p, sizeof(float2) * n, hipMemcpyHostToDevice), "uploadDensity");
}

int GPUSPHFluid::writePly(const std::string& path) const {
    const int n = mFluidParams.max_particles;
    std::vector<float4> pos(n > 0 ? n : 0), vel(pos.size());
    std::vector<float2> rp(pos.size());
    std::vector<unsigned int> id(pos.size());
    if (n > 0) {
        hip_check(hipDeviceSynchronize(), "writePly sync");
        hip_check(hipMemcpy(pos.data(), d_position, sizeof(float4) * n, hipMemcpyDeviceToHost), "writePly");
        hip_check(hipMemcpy(vel.data(), d_velocity, sizeof(float4) * n, hipMemcpyDeviceToHost), "writePly");
        hip_check(hipMemcpy(rp.data(), d_mass_pressure, sizeof(float2) * n, hipMemcpyDeviceToHost), "writePly");
        hip_check(hipMemcpy(id.data(), d_particleId, sizeof(unsigned int) * n, hipMemcpyDeviceToHost), "writePly");
    }
    // (rho, p) are stored in sorted order: slot k belongs to particle id[k].
    std::vector<float2> byParticle(pos.size(), make_float2(0.0f, 0.0f));
    for (int k = 0; k < n; k++)
        if (id[k] < (unsigned)n) byParticle[id[k]] = rp[k];
    int count = 0;
    for (int i = 0; i < n; i++) count += pos[i].w > 0.0f;
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("GPUSPHFluid::writePly: cannot open " + path);
    std::fprintf(f,
                 "ply\nformat binary_little_endian 1.0\ncomment fluidgrid_amd particles\nelement vertex %d\n"
                 "property float x\nproperty float y\nproperty float z\nproperty float vx\nproperty float vy\n"
                 "property float vz\nproperty float density\nproperty float pressure\nend_header\n",
                 count);
    bool ok = true;
    for (int i = 0; i < n && ok; i++) {
        if (!(pos[i].w > 0.0f)) continue;
        const float rec[8] = {pos[i].x, pos[i].y, pos[i].z, vel[i].x, vel[i].y, vel[i].z, byParticle[i].x,
                              byParticle[i].y};
        ok = std::fwrite(rec, sizeof(rec), 1, f) == 1;
    }
    if (std::fclose(f) != 0 || !ok) throw std::runtime_error("GPUSPHFluid::writePly: write failed for " + path);
    return count;
}

void GPUSPHFluid::emitInflow(float dt) {
    if (d_emitStep)
        fg_emit_inflow(reinterpret_cast<fg_float4*>(d_position), reinterpret_cast<fg_float4*>(d_velocity), d_emitStep,
                       &mInflow, dt);
}

}  // namespace Fluid
}  // namespace Simulation
