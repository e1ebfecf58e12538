/*
 * fluidgrid_amd_sim.h -- C-ABI over the C++ HeterogeneousSimulator mirror
 * (include/fluidgrid/HeterogeneousSimulator.h).  No counterpart in the
 * reference: it lets callers without a C++ toolchain (FFI, Python ctypes)
 * drive the whole step the way the reference's application does
 * (AppMain.cpp:78-133 then HeterogeneousSimulator::doTimeStep per frame).
 * One process-wide simulator; not reentrant.
 */
#ifndef FLUIDGRID_AMD_SIM_H
#define FLUIDGRID_AMD_SIM_H

#include "fluidgrid_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fg_fluid_buffers {
    void* hashKey;          /* u32[n]      sorted cell keys            */
    void* particleId;       /* u32[n]      sorted particle ids         */
    void* gridStart;        /* u32[nb]     first sorted index per cell */
    void* gridEnd;          /* u32[nb]     one past the last           */
    void* position;         /* float4[n]   particle order, w = active  */
    void* orderedPosition;  /* float4[n]   sorted order                */
    void* velocity;         /* float4[n]   particle order              */
    void* orderedVelocity;  /* float4[n]   sorted order                */
    void* densityPressure;  /* float2[n]   (rho, p), sorted order      */
    void* forces;           /* float4[n]   unused by the step          */
    unsigned int numParticles;
    unsigned int numBuckets;
} fg_fluid_buffers;

/* Resets the simulator and makes `common` the parameter block
 * (ApplicationData::setCommonFluidParams, then the simulator constructor). */
int fg_sim_init(const fg_common_fluid_params* common, int key_mode, int math_mode);
/* Adds a fluid with host-side initial state; returns its index, or a
 * negated FG_ERR_* code. */
int fg_sim_add_fluid(const fg_fluid_params* params, const fg_float4* positions, const fg_float4* velocities);
int fg_sim_set_options(int fused, int use_graph, int stage_timing);
/* With rigid bodies: run the body work on a second stream beside the fluid
 * stages (default on; results are the same either way). */
int fg_sim_set_overlap(int overlap_bodies);
/* Density arithmetic of the fused step: FG_PRECISION_FP32 (default) or
 * FG_PRECISION_MIXED (fluidgrid_amd.h). */
int fg_sim_set_precision(int mode);
/* Launch-shape tuning of the single-fluid step (process-wide:
 * fg_set_tuning, fg_set_schedule, fg_set_tiled): times each candidate for
 * trialSteps graph steps from the same state, keeps the fastest and restores
 * the simulation state.  tuning (optional, 5 ints) receives (sphBlock,
 * sortDigitBits, flatDensity, flatForces, tiled) of the choice; trials
 * (optional, maxTrials records of 6 floats) the same five and the ms per
 * step of each candidate timed; *numTrials their count.  No choice changes
 * any result bit. */
int fg_sim_autotune(int trialSteps, int* tuning, float* trials, int maxTrials, int* numTrials);
int fg_sim_step(int steps);                       /* asynchronous */
int fg_sim_synchronize(void);
void* fg_sim_stream(void);                        /* hipStream_t of the simulator */
int fg_sim_download(int fluid, fg_float4* positions, fg_float4* velocities);
int fg_sim_upload(int fluid, const fg_float4* positions, const fg_float4* velocities);
int fg_sim_fluid_buffers(int fluid, fg_fluid_buffers* out);
/* hash, sort, cells, density, forces, integrate, bodies (ms) of the last step
 * run with stage_timing on (7 floats). */
int fg_sim_stage_times(float* out7);

/* Attaches an inflow to fluid idx: its reserve slots (inactive, parked away
 * from the fluid) are emitted layer by layer at the start of every step from
 * now on (fg_emit_inflow; the emission step counter starts at 0). */
int fg_sim_set_inflow(int idx, const fg_inflow_desc* desc);

/* FP64 conservation sums of fluid idx (fg_conservation_f64) into host out8:
 * count, mass, momentum x/y/z, kinetic, gravitational potential and EOS
 * internal energy.  Synchronises the simulator's stream. */
int fg_sim_conservation(int idx, double* out8);

/* Dynamic rigid bodies coupled to the (single) fluid: sphere centres in
 * world space, body b owning [bodyOffsets[b], bodyOffsets[b+1]); masses in
 * kg; the coupling parameters (ApplicationData::setCommonRigidBodyParams /
 * setSimulationParams).  Centre of mass and inertia are computed on the GPU
 * in FP64 from the spheres. */
int fg_sim_add_bodies(const fg_float4* worldSpheres, const unsigned int* bodyOffsets, unsigned int numBodies,
                      const double* masses, float sphereRadius, const fg_common_rigid_body_params* rigidParams,
                      float springForce, float dampingForce);
/* Static rigid spheres (one static body, RigidBody::isStatic; the static
 * grid of RigidBodySimulator.cpp:146-214) that the fluid collides with:
 * keyed with the reference hash over st_bucket_size buckets of the rigid
 * cell size, sorted once here.  Every call adds a body; call before the
 * first step (the rigid-body state is rebuilt). */
int fg_sim_add_static_spheres(const fg_float4* worldSpheres, unsigned int numSpheres, float sphereRadius,
                              const fg_common_rigid_body_params* rigidParams, float springForce,
                              float dampingForce);
/* The static grid built by fg_sim_add_static_spheres (device pointers owned
 * by the simulator; numSpheres is 0 when there is none). */
int fg_sim_static_grid(fg_body_grid* out);
/* Per-body centre-of-mass position, velocity, orientation (x,y,z,w) and
 * angular momentum; any pointer may be NULL. */
int fg_sim_body_state(fg_float4* positions, fg_float4* velocities, fg_float4* orientations,
                      fg_float4* angularMomenta);
/* FP64 force and torque sums of the last step (4 doubles per body, w = 0). */
int fg_sim_body_sums(double* forces4, double* torques4);
/* FP64 centre of mass (3 per body) and inertia tensor (9 per body). */
int fg_sim_body_mass_properties(double* centreOfMass3, double* inertia9);

/* Checkpoint / resume (HeterogeneousSimulator::saveState / loadState; the
 * reference has none).  save writes the state a step reads from the step
 * before (particles, the sorted density buffer, inflow step counters, body
 * state) with the parameter blocks and a checksum.  load restores it into a
 * simulator set up with the same scene; stepping on from there gives the
 * same bits as an uninterrupted run.  load changes nothing and returns
 * FG_ERR_INVALID_ARGUMENT if the file does not match the scene or fails its
 * checksum.  Both synchronise the device. */
int fg_sim_save_state(const char* path);
int fg_sim_load_state(const char* path);
/* Point cloud of fluid idx for visualisation (GPUSPHFluid::writePly): a
 * binary little-endian PLY file of the active particles with position,
 * velocity, density and pressure.  *count (may be NULL) receives the number
 * of vertices written. */
int fg_sim_write_ply(int idx, const char* path, unsigned int* count);
int fg_sim_shutdown(void);

#ifdef __cplusplus
}
#endif

#endif
