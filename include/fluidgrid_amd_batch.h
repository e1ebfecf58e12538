/* fluidgrid_amd_batch.h -- C-ABI over the batched-scene simulator
 * (include/fluidgrid/BatchedSimulator.h).  No counterpart in the reference:
 * it advances many independent fluids, each with its own parameter block,
 * with one launch per stage.  One process-wide batch. */
#ifndef FLUIDGRID_AMD_BATCH_H
#define FLUIDGRID_AMD_BATCH_H

#include "fluidgrid_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Creates the batch (replacing any previous one).  Scene s has counts[s]
 * particles; positions/velocities hold all scenes back to back (scene-major).
 * Every commons[s] describes a LINEAR grid (bucket_size, cell_size,
 * volume_min) and the scene's box, gravity, dt and scale. */
int fg_bsim_create(unsigned int numScenes, const fg_common_fluid_params* commons, const fg_fluid_params* fluids,
                   const unsigned int* counts, const fg_float4* positions, const fg_float4* velocities);
/* Dynamic rigid bodies of one scene (the arguments of fg_sim_add_bodies),
 * coupled two-way with that scene's fluid by the stages and arithmetic of
 * the single-fluid coupled step: each scene with bodies advances exactly as
 * a single-fluid simulator given the same fluid and bodies.  Bodies of
 * different scenes never interact.  Before the first step, once per scene. */
int fg_bsim_add_bodies(unsigned int scene, const fg_float4* worldSpheres, const unsigned int* bodyOffsets,
                       unsigned int numBodies, const double* masses, float sphereRadius,
                       const fg_common_rigid_body_params* rigidParams, float springForce, float dampingForce);
/* Body count (numBodies, may be NULL) and, for every body with the scenes in
 * order (the bodies of the first scene given bodies first), centre of mass,
 * velocity, orientation (x, y, z, w) and angular momentum; arrays may be NULL.
 * Bodies become part of the batch at its first step: before it the count is 0. */
int fg_bsim_body_state(unsigned int* numBodies, fg_float4* positions, fg_float4* velocities, fg_float4* orientations,
                       fg_float4* angularMomenta);
int fg_bsim_set_options(int use_graph, int stage_timing);
/* Body stages on a second stream beside the fluid stages (default 1; the
 * results are the same bits either way; stage timing runs them in line). */
int fg_bsim_set_overlap(int overlap_bodies);
int fg_bsim_step(int steps);                      /* asynchronous */
int fg_bsim_synchronize(void);
void* fg_bsim_stream(void);                       /* hipStream_t of the batch (null without one) */
int fg_bsim_download(fg_float4* positions, fg_float4* velocities);
/* (max(1,rho), p) in sorted order and the sorted particle ids of the last step. */
int fg_bsim_download_density(fg_float2* rho_p, unsigned int* particleId);
/* hash, sort, cells, density, forces+integrate (ms) of the last timed step. */
int fg_bsim_stage_times(float* out5);
/* numParticles, numBuckets, keyBits of the batch. */
int fg_bsim_info(unsigned int* numParticles, unsigned int* numBuckets, int* keyBits);
/* Launch-shape tuning: times each candidate for trialSteps steps, keeps the
 * fastest and restores the particle state.  tuning (optional, 6 ints)
 * receives (sphBlock, sortDigitBits, flatDensity, flatForces, queueItems,
 * deferContact) of the choice; trials (optional, capacity maxTrials records
 * of 7 floats) receives one (sphBlock, sortDigitBits, flatDensity,
 * flatForces, queueItems, deferContact, ms per step) record per candidate
 * timed; *numTrials their count. */
int fg_bsim_autotune(int trialSteps, int* tuning, float* trials, int maxTrials, int* numTrials);
/* queueItems: 0 launches one workgroup per block of particles; 1..64 runs
 * the density and force passes as persistent launches (as many workgroups
 * as the GPU holds at once) that take work items of that many blocks'
 * particles from per-XCD counters.  deferContact: with bodies, the coupled
 * force pass runs as a force kernel without the body-contact walk, which
 * finishes the particles outside every body grid's bounds, and a second
 * pass that finishes the others with it (default 1).  No choice changes any
 * result bit. */
int fg_bsim_set_tuning(int sphBlock, int sortDigitBits, int flatDensity, int flatForces, int queueItems,
                       int deferContact);
int fg_bsim_destroy(void);

#ifdef __cplusplus
}
#endif

#endif /* FLUIDGRID_AMD_BATCH_H */
