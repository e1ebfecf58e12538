// reference_caller.cpp -- a caller built the way the reference's host code
// is: HIP's own vector types (float2/float3/float4/int3) and bool passed by
// value, the parameter block declared as the reference declares it
// (ApplicationData.cuh:46-67), the launchers declared with the reference's
// prototypes (K_FluidSim.cuh:43-44, K_HashGrid.cuh:5-8, K_FluidSim.cuh:46-78)
// and none of this library's headers.  It links against
// libfluidgrid_amd.so and prints what the library received, so a test can
// check that every by-value argument arrived intact.  It makes no device
// call: every launcher below stops at its argument check.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

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
};

extern "C" {
void gInitializeCommonFluidParams(CommonFluidParams params);
void gCalcHashKey(unsigned int* hashKey, unsigned int* particleID, float4* positions, unsigned int numParticles,
                  unsigned int numBuckets, float3 cell_size);
void gSPHCalculateForcesWithSRB(float4* oldPositions, float4* oldVelocity, float4* forces, float4* newVel,
                                float2* density_pressure, unsigned int* particleId, unsigned int* gridStart,
                                unsigned int* gridEnd, unsigned int numParticles, float mass, float kViscosity,
                                float sph_particle_radius, int3 cell_size_relation, float srb_particle_radius,
                                float3 srb_cell_size, float kspring, float kdamping, float4* srb_positions,
                                unsigned int* srb_start_offset, unsigned int* srb_end_offset,
                                unsigned int srb_num_buckets, float4* rb_positions, float4* rb_velocities,
                                unsigned int* rb_start_offset, unsigned int* rb_end_offset,
                                unsigned int rb_num_buckets, bool useStaticRB, bool useDynamicRB);
// extensions used to read back what arrived
int fg_get_common_fluid_params(CommonFluidParams* out);
int fg_last_error(void);
const char* fg_last_error_string(void);
}

int main() {
    CommonFluidParams cp;
    std::memset(&cp, 0, sizeof(cp));
    cp.cell_size = make_float3(0.015625f, 0.03125f, 0.0625f);
    cp.bucket_size = make_int3(17, 19, 23);
    cp.keyBits = 13;
    cp.kernel_radius = 0.125f;
    cp.kernel_radius_squared = 0.015625f;
    cp.pointRadius = 0.25f;
    cp.pointScale = 3.0f;
    cp.velocity_limit = 600.0f;
    cp.volume_min = make_float3(-1.5f, -2.5f, -3.5f);
    cp.volume_max = make_float3(4.5f, 5.5f, 6.5f);
    cp.global_external_force = make_float3(0.0f, -9.75f, 0.5f);
    cp.simulation_scale = 0.0078125f;
    cp.deltaTimeIntegration = 0.00390625f;
    cp.fluid_to_srb_relation = make_int3(1, 2, 3);
    gInitializeCommonFluidParams(cp);
    CommonFluidParams back;
    std::memset(&back, 0xCD, sizeof(back));
    const int rc = fg_get_common_fluid_params(&back);
    std::printf("sizeof %zu\n", sizeof(CommonFluidParams));
    std::printf("common_roundtrip rc=%d equal=%d\n", rc, std::memcmp(&cp, &back, sizeof(cp)) == 0 ? 1 : 0);

    // Non-null stand-ins: the argument checks below run before any use.
    float4* p4 = reinterpret_cast<float4*>(256);
    float2* p2 = reinterpret_cast<float2*>(256);
    unsigned int* pu = reinterpret_cast<unsigned int*>(256);

    gCalcHashKey(pu, pu, p4, 4321u, 98765u, make_float3(-0.5f, 0.25f, 1.5f));
    std::printf("hash err=%d msg=%s\n", fg_last_error(), fg_last_error_string());

    gSPHCalculateForcesWithSRB(p4, p4, nullptr, p4, p2, pu, pu, pu, 12345u, 0.25f, 3.5f, 0.125f, make_int3(1, 2, 3),
                               0.0625f, make_float3(0.5f, 0.75f, 1.25f), -2000.0f, -50.0f, nullptr, nullptr, nullptr,
                               4096u, p4, p4, pu, pu, 777u, true, false);
    std::printf("srb err=%d msg=%s\n", fg_last_error(), fg_last_error_string());
    return 0;
}
