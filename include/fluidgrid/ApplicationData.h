// ApplicationData.h -- parameter model of the SPH step, member for member the
// layout of the reference's ApplicationData.cuh:23-159 (same names, same
// field order, HIP vector types in place of CUDA's), so host code written
// against the reference compiles unchanged with hipcc.
#pragma once

#include <hip/hip_runtime.h>

#include "../fluidgrid_amd.h"

namespace Simulation {
namespace Application {

static const float PI = 3.14159265f;

// ApplicationData.cuh:23-27
struct SimulationParams {
    float dampingForce;
    float springForce;
};

// ApplicationData.cuh:32-41
struct FluidParams {
    float mass;
    float restDensity;
    float viscosity;
    float stiffness;
    float particle_radius;
    float gasConstant;
    int max_particles;
};

// ApplicationData.cuh:46-67
struct CommonFluidParams {
    float3 cell_size;
    int3 bucket_size;
    unsigned int keyBits;
    float kernel_radius;
    float kernel_radius_squared;
    float pointRadius;
    float pointScale;
    float velocity_limit;
    float3 volume_min;
    float3 volume_max;
    float3 global_external_force;
    float simulation_scale;
    float deltaTimeIntegration;
    int3 fluid_to_srb_relation;

    __host__ __device__ int getNumBuckets() const { return bucket_size.x * bucket_size.y * bucket_size.z; }
};

// ApplicationData.cuh:73-102
struct CommonRigidBodyParams {
    int3 rb_to_fluid_relation;
    int3 st_bucket_size;
    int3 dn_bucket_size;
    unsigned int st_keyBits;
    float velocity_limit;
    float3 volume_min;
    float3 volume_max;
    float4 global_external_force;
    float simulation_scale;
    float deltaTimeIntegration;
    float particleRadius;

    __host__ __device__ int getStNumBuckets() const {
        return st_bucket_size.x * st_bucket_size.y * st_bucket_size.z;
    }
    __host__ __device__ int getDnNumBuckets() const {
        return dn_bucket_size.x * dn_bucket_size.y * dn_bucket_size.z;
    }
    __host__ __device__ float3 getCellSize() const {
        return make_float3(particleRadius, particleRadius, particleRadius);
    }
};

// ApplicationData.cuh:107-112
struct KernelsCache {
    float kPoly6;
    float kSpikyGradient;
    float kViscosityLaplacian;
};

static_assert(sizeof(FluidParams) == sizeof(fg_fluid_params), "FluidParams layout");
static_assert(sizeof(CommonFluidParams) == sizeof(fg_common_fluid_params), "CommonFluidParams layout");
static_assert(sizeof(CommonRigidBodyParams) == sizeof(fg_common_rigid_body_params), "CommonRigidBodyParams layout");
static_assert(sizeof(KernelsCache) == sizeof(fg_kernels_cache), "KernelsCache layout");

// ApplicationData.cuh:118-159 (singleton holding the parameter blocks)
class ApplicationData {
public:
    static ApplicationData* instance();

    CommonFluidParams& getCommonFluidParams() { return mCommonFluidParams; }
    KernelsCache& getKernelsCache() { return mKernelsCache; }
    CommonRigidBodyParams& getCommonRigidBodyParams() { return mCommonRigidBodyParams; }
    SimulationParams& getSimulationParams() { return mSimulationParams; }

    void setSimulationParams(SimulationParams p) { mSimulationParams = p; }
    void setCommonFluidParams(CommonFluidParams p) { mCommonFluidParams = p; }
    void setCommonRigidBodyParams(CommonRigidBodyParams p) { mCommonRigidBodyParams = p; }
    void setKernelsCache(KernelsCache k) { mKernelsCache = k; }

private:
    ApplicationData() = default;
    CommonFluidParams mCommonFluidParams{};
    CommonRigidBodyParams mCommonRigidBodyParams{};
    SimulationParams mSimulationParams{};
    KernelsCache mKernelsCache{};
};

// Bit-copies between the C++ parameter structs and the C-ABI ones.
template <class To, class From>
inline To abi_cast(const From& f) {
    static_assert(sizeof(To) == sizeof(From), "layout mismatch");
    To t;
    __builtin_memcpy(&t, &f, sizeof(To));
    return t;
}

}  // namespace Application
}  // namespace Simulation
