/*
 * fluidgrid_amd.h -- C-ABI of the MI355X (gfx950) per-timestep SPH pipeline.
 *
 * Two groups of entry points live here:
 *
 *  1. The reference launchers, same names, same argument order and meaning,
 *     same `void` return.  Each declaration cites the reference header line it
 *     replaces (paths relative to the FluidGrid source tree).  A host program
 *     written against GPU_Kernels/K_HashGrid.cuh, K_FluidSim.cuh and
 *     K_RigidBodySim.cuh links against libfluidgrid_amd.so unchanged.
 *
 *  2. `fg_*` extensions: stream selection, error query, key/math modes, an
 *     explicit-key-width sort, the fused force+integrate launcher, device
 *     memory helpers for callers that do not link a HIP runtime themselves,
 *     per-body FP64 reductions, and the opaque simulator handle used by the
 *     C++ `HeterogeneousSimulator` mirror.
 *
 * Vector types.  fg_float2/3/4 and fg_int3 are plain structs with the size
 * and alignment of the HIP/CUDA vector types (8/8, 12/4, 16/16, 12/4); they
 * are passed by value exactly like HIP's float3/float4 on x86-64 SysV.
 *
 * Conventions shared by every launcher:
 *  - all pointers are device pointers owned by the caller; launchers never
 *    allocate, except gSort/fg_sort_pairs which keep a persistent scratch
 *    buffer that grows on demand (fg_sort_pairs_ws takes caller workspace);
 *  - work is enqueued on the stream chosen by fg_set_stream (default: the
 *    null stream) and launchers do not synchronise;
 *  - launchers validate arguments and record failures in a thread-local
 *    error slot read by fg_last_error(); they never abort the process.
 */
#ifndef FLUIDGRID_AMD_H
#define FLUIDGRID_AMD_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fg_float2 { float x, y; } __attribute__((aligned(8))) fg_float2;
typedef struct fg_float3 { float x, y, z; } fg_float3;
typedef struct fg_float4 { float x, y, z, w; } __attribute__((aligned(16))) fg_float4;
typedef struct fg_int3 { int x, y, z; } fg_int3;
typedef struct fg_double4 { double x, y, z, w; } fg_double4;

/* ApplicationData.cuh:32-41 (28 bytes) */
typedef struct fg_fluid_params {
    float mass;
    float restDensity;
    float viscosity;
    float stiffness;
    float particle_radius;
    float gasConstant;
    int   max_particles;
} fg_fluid_params;

/* ApplicationData.cuh:46-67 (104 bytes, align 4) */
typedef struct fg_common_fluid_params {
    fg_float3    cell_size;
    fg_int3      bucket_size;
    unsigned int keyBits;
    float        kernel_radius;
    float        kernel_radius_squared;
    float        pointRadius;
    float        pointScale;
    float        velocity_limit;
    fg_float3    volume_min;
    fg_float3    volume_max;
    fg_float3    global_external_force;
    float        simulation_scale;
    float        deltaTimeIntegration;
    fg_int3      fluid_to_srb_relation;
} fg_common_fluid_params;

/* ApplicationData.cuh:73-102 (112 bytes, align 16) */
typedef struct fg_common_rigid_body_params {
    fg_int3      rb_to_fluid_relation;
    fg_int3      st_bucket_size;
    fg_int3      dn_bucket_size;
    unsigned int st_keyBits;
    float        velocity_limit;
    fg_float3    volume_min;
    fg_float3    volume_max;
    fg_float4    global_external_force;
    float        simulation_scale;
    float        deltaTimeIntegration;
    float        particleRadius;
} fg_common_rigid_body_params;

/* ApplicationData.cuh:107-112 (12 bytes) */
typedef struct fg_kernels_cache {
    float kPoly6;
    float kSpikyGradient;
    float kViscosityLaplacian;
} fg_kernels_cache;

/* ------------------------------------------------------------------------ */
/* 1. Reference launchers                                                    */
/* ------------------------------------------------------------------------ */

/* K_FluidSim.cuh:43-44 -- makes `p` the parameter block of every later
 * launch (held host-side and passed as a kernel argument). */
void gInitializeCommonFluidParams(fg_common_fluid_params p);

/* K_FluidSim.cuh:18-19 -- kPoly6, kSpikyGradient, kViscosityLaplacian. */
void gInitializeKernelParams(fg_kernels_cache k);

/* K_HashGrid.cuh:5-8 -- key[i] = cell key of positions[i]; particleID[i] = i.
 * The key function is the one selected by fg_set_key_mode. */
void gCalcHashKey(unsigned int* hashKey, unsigned int* particleID, fg_float4* positions,
                  unsigned int numParticles, unsigned int numBuckets, fg_float3 cell_size);

/* K_HashGrid.cuh:10-20 -- body-sphere world transform conj(q)*p*q + cm[w],
 * then key and identity id. */
void gCalcHashKeyWithOffset(unsigned int* _out_hashKey, unsigned int* _out_particleID,
                            fg_float4* _in_vectorOffset, fg_float4* _in_orientation,
                            fg_float4* _in_vectorPos, fg_float4* _out_positions,
                            unsigned int _in_numParticles, fg_float3 _in_cellSize,
                            int _in_hashTableSize);

/* K_FluidSim.cuh:81-82 -- stable in-place key/value sort (ties keep input
 * order).  The key width is found on the device; one 4-byte readback. */
void gSort(unsigned int* _key, unsigned int* _values, int _numParticles);

/* K_HashGrid.cuh:23-34 -- cellStart := 0xFFFFFFFF over numBuckets; cell
 * start/end from key transitions of the sorted keys; SoA gather of pos/vel
 * through the sorted ids.  cellEnd is not cleared. */
void gFindCellOffsetOrderingPosVel(unsigned int* gridPosStart, unsigned int* gridPosEnd,
                                   unsigned int* hashTable, unsigned int* particleID,
                                   fg_float4* particlePositions, fg_float4* orderedParticlePositions,
                                   fg_float4* particleVel, fg_float4* particleOrderedVel,
                                   unsigned int numParticles, unsigned int numBuckets);

/* K_HashGrid.cuh:37-56 -- as above, positions only. */
void gFindCellOffsetOrderingPos(unsigned int* gridPosStart, unsigned int* gridPosEnd,
                                unsigned int* hashTable, unsigned int* particleID,
                                fg_float4* particlePositions, fg_float4* orderedParticlePositions,
                                unsigned int numParticles, unsigned int numBuckets);

/* K_FluidSim.cuh:7-16 -- (max(1,rho), k*(rho-rho0)) for active particles,
 * poly6 over the 27-cell stencil; hashTable is accepted and unused. */
void gCalcMassDensity(fg_float4* orderedPositions, fg_float2* mass_pressure,
                      unsigned int* hashTable, unsigned int* gridStart, unsigned int* gridEnd,
                      int numParticles, float gasConstant, float restDensity, float mass);

/* K_FluidSim.cuh:21-32 -- spiky pressure + Laplacian viscosity + gravity +
 * penalty walls + clamp; vel[particleId[i]] = orderedVelocity[i] + a*dt.
 * `forces` is accepted and unused. */
void gSPHCalculateForces(fg_float4* orderedPositions, fg_float4* orderedVelocity,
                         fg_float4* forces, fg_float2* density_pressure,
                         unsigned int* gridStart, unsigned int* gridEnd, fg_float4* vel,
                         unsigned int* particleId, unsigned int numParticles,
                         float mass, float viscosity);

/* K_FluidSim.cuh:34-41 -- pos += vel*(dt/scale) where pos.w > 0.
 * forces/density/grid arguments are accepted and unused. */
void gSPHIntegrate(fg_float4* _in_sph_positions, fg_float4* _in_sph_velocity,
                   fg_float4* _in_sph_forces, unsigned int _in_num_particles,
                   fg_float2* _in_density_press, unsigned int* _in_grid_start,
                   unsigned int* _in_grid_end);

/* K_FluidSim.cuh:46-78 -- gSPHCalculateForces plus penalty contact against a
 * static body grid and a dynamic body grid (velocities indexed by body id). */
void gSPHCalculateForcesWithSRB(fg_float4* oldPositions, fg_float4* oldVelocity,
                                fg_float4* forces, fg_float4* newVel,
                                fg_float2* density_pressure, unsigned int* particleId,
                                unsigned int* gridStart, unsigned int* gridEnd,
                                unsigned int numParticles, float mass, float kViscosity,
                                float _in_sph_particle_radius, fg_int3 _in_cell_size_relation,
                                float _in_srb_particle_radius, fg_float3 _in_srb_cell_size,
                                float _in_kspring, float _in_kdamping,
                                fg_float4* _in_srb_positions, unsigned int* _in_srb_start_offset,
                                unsigned int* _in_srb_end_offset, unsigned int _in_srb_num_buckets,
                                fg_float4* _in_rb_positions, fg_float4* _in_rb_velocities,
                                unsigned int* _in_rb_start_offset, unsigned int* _in_rb_end_offset,
                                unsigned int _in_rb_num_buckets,
                                bool useStaticRB, bool useDynamicRB);

/* K_RigidBodySim.cuh:5-25 -- per body sphere: penalty force from the spheres
 * of the other dynamic bodies (same-body spheres skipped by w, body
 * velocities indexed by w) and, when _in_numSRB > 0, from the static spheres;
 * contact within 2*_in_Radius + 0.001, stencil (1,1,1).  _in_RBPositions is
 * both the query set and the sorted grid array.  Writes F and the torque
 * (x - cm[w]) x F, both with w = body id. */
void gCalculateRigidBodyForcesFromRB(fg_float4* _in_RBPositions, fg_float4* _in_RBVelocites,
                                     fg_float4* _in_RBCMPosition, fg_float4* _out_torque, fg_float4* _out_Forces,
                                     unsigned int* _in_RBGridStart, unsigned int* _in_RBGridEnd,
                                     unsigned int _in_NumParticles, fg_float3 _in_CellSize, float _in_Radius,
                                     int _in_RBHashTableSize, int _in_SRBHashTableSize, fg_float4* _in_SRBPositions,
                                     unsigned int* _in_SRBGridStart, unsigned int* _in_SRBGridEnd,
                                     float _in_RBkSpring, float _in_RBkDamping, int _in_numSRB);

/* K_RigidBodySim.cuh:27-50 -- per body sphere: penalty force from the fluid
 * grid and torque (x - cm[w]) x F; F.w = torque.w = body id. */
void gCalculateRigidBodyForcesFromFluids(fg_float4* _in_RBPositions, fg_float4* _in_RBVelocites,
                                         fg_float4* _in_RBCMPos, fg_float4* _out_Forces,
                                         fg_float4* _out_torque, unsigned int* _in_RBGridStart,
                                         unsigned int* _in_RBGridEnd, unsigned int _in_NumParticles,
                                         fg_float3 _in_CellSize, float _in_Radius,
                                         int _in_RBHashTableSize, fg_float4* _in_sph_particles,
                                         fg_float4* _in_sph_velocities,
                                         unsigned int* _in_sph_grid_start,
                                         unsigned int* _in_sph_grid_end,
                                         float _in_sph_particle_radius, fg_float3 _in_sph_cell_size,
                                         unsigned int _in_sph_num_buckets,
                                         fg_int3 _in_cell_size_relation,
                                         float _in_sph_kSpring, float _in_sph_kDamping);

/* K_RigidBodySim.cuh:52-55 -- key = (uint)F.w (body id), value = index. */
void gCopyRBDataToParticle(unsigned int* _out_force_rb, unsigned int* _out_force_index,
                           fg_float4* _in_rb_particle_forces, size_t num_particles);

/* K_RigidBodySim.cuh:57-59 -- gather F and torque through sorted indices. */
void gReorderForceTorque(fg_float4* _out_force, fg_float4* _out_torque, fg_float4* _in_force,
                         fg_float4* _in_torque, unsigned int* _in_ordered_index,
                         unsigned int _num_values);

/* K_RigidBodySim.cuh:61-62 -- body[F.w] += scan value at each segment end. */
void gAddForceTorque(fg_float4* _in_sum_force, fg_float4* _in_sum_torque,
                     fg_float4* _in_ordened_force, fg_float4* _out_to_force,
                     fg_float4* _out_to_torque, int* _in_sum_offset, size_t num_rigid_body);

/* K_RigidBodySim.cuh:63-65 -- body linear/angular integration with floor and
 * side-wall bounce. */
void gIntegrateRB(fg_float4* _in_RBPosition, float* _in_mass, fg_float4* _in_RBVelocities,
                  fg_float4* _in_rbforces, size_t num_rb, float _elapsed_time_in_seconds,
                  fg_float4 _in_external_force, fg_float3 min_volume, fg_float3 max_volume,
                  fg_float4* _angular_momentum, fg_float4* _torque, fg_float4* _inertia_tensor,
                  fg_float4* _orientation, fg_float4* _inv_inertia_tensor, float velocity_limit);

/* ------------------------------------------------------------------------ */
/* 2. Extensions (no counterpart in the reference)                           */
/* ------------------------------------------------------------------------ */

enum {
    FG_OK = 0,
    FG_ERR_INVALID_ARGUMENT = 1,
    FG_ERR_HIP = 2,
    FG_ERR_OUT_OF_MEMORY = 3,
    FG_ERR_NOT_INITIALIZED = 4,
    FG_ERR_UNSUPPORTED = 5
};

/* Cell key function used by gCalcHashKey and by the neighbour stencil of
 * the density/force kernels.
 *  FG_KEY_REF_XOR: ((xi*73856093)^(yi*19349663)^(zi*83492791)) % numBuckets
 *    with xi = (int)((p.x+5000)/cs.x)  -- K_DeviceFunctions.cu:17-26;
 *    the stencil rehashes p + k*cs (K_FluidSim.cu:185-203).
 *  FG_KEY_LINEAR:  key = cx + nx*(cy + ny*cz), ci = clamp(floor((p-vmin)/cs),
 *    0, n-1) with n = bucket_size; the stencil is ci-1..ci+1, clipped. */
enum { FG_KEY_REF_XOR = 0, FG_KEY_LINEAR = 1 };

/* Arithmetic of the density/force kernels.
 *  FG_MATH_FAST: neighbour rows walked as contiguous sorted ranges, cell
 *    culling, common factors hoisted out of the pair loop (default).
 *  FG_MATH_REFERENCE: per-bucket partial sums and the reference's per-pair
 *    expressions, no contraction. */
enum { FG_MATH_FAST = 0, FG_MATH_REFERENCE = 1 };

/* Arithmetic of the fused (masked, FG_MATH_FAST, FG_KEY_LINEAR) density pass.
 *  FG_PRECISION_FP32: every candidate tested and summed in FP32 (default).
 *  FG_PRECISION_MIXED: the cell pass also writes each sorted particle's
 *    offset inside its cell as FP16 (fg_cell_offsets_pos_vel with a
 *    neighbour buffer); density culls candidates two at a time in packed
 *    FP16 with a margin that never drops a pair the FP32 test accepts,
 *    evaluates the survivors exactly as FP32 does, and sums them with
 *    Neumaier compensation.  The row masks and neighbour records are those
 *    of FP32, so the force pass is unchanged.  Mode of this process; set it
 *    between steps (a density pass in the mixed mode needs the FP16 offsets
 *    of the same sorted order, and refuses a buffer without them). */
enum { FG_PRECISION_FP32 = 0, FG_PRECISION_MIXED = 1 };

int         fg_last_error(void);          /* last error code of this thread, then FG_OK */
const char* fg_last_error_string(void);   /* message of the last error of this thread */
int         fg_peek_error(void);          /* last error code without clearing */
void        fg_record_error(int code, const char* message); /* sets this thread's error slot */
const char* fg_version(void);

int   fg_set_stream(void* hip_stream);    /* hipStream_t; NULL = null stream */
void* fg_get_stream(void);
int   fg_set_key_mode(int mode);
int   fg_get_key_mode(void);
int   fg_set_math_mode(int mode);
int   fg_get_math_mode(void);
/* Launch shape (process-wide): workgroup size of the fast density/force
 * kernels (64, 128, 256 or 512; default 64) and the widest radix digit of
 * the sort (6..8 bits; default 8).  Results do not depend on either. */
int   fg_set_tuning(int sphBlock, int sortDigitBits);
int   fg_get_tuning(int* sphBlock, int* sortDigitBits);
/* Schedules of the masked fast density and force passes: 1 walks each
 * particle's candidates (neighbours) as one stream across the 9 stencil rows
 * with a per-lane row table in LDS, 0 (default) row by row.  The summation
 * order, and so every result bit, is the same. */
int   fg_set_schedule(int flatDensity, int flatForces);
int   fg_get_schedule(int* flatDensity, int* flatForces);
/* LDS-tiled masked passes (default off): fg_sph_density_nbr stages each
 * 128-particle tile's stencil rows in LDS and writes a neighbour list per
 * particle; the force pass that reads the buffer walks the lists over
 * LDS-staged neighbour records.  They take precedence over the schedules
 * above and give the same bits.  The force pass follows the form the last
 * density pass wrote into its neighbour buffer, whatever is set when it
 * runs. */
int   fg_set_tiled(int enable);
int   fg_get_tiled(int* enable);
/* Coupled force pass with neighbour masks, row-by-row schedule (default on):
 * the particles inside a body grid's bounds hand their pair sum to a second
 * kernel that adds body contact (walking each one's body-grid cells on two
 * lanes), walls and clamp, and integrates them; the others finish in the
 * force pass without the contact walk.  No counterpart in the reference: a
 * launch-shape knob like fg_set_schedule; every result bit is the same.  It
 * uses the neighbour buffer's list area, which only the LDS-tiled passes
 * use otherwise. */
int   fg_set_contact_defer(int enable);
int   fg_get_contact_defer(int* enable);
int   fg_set_precision(int mode);
int   fg_get_precision(void);
int   fg_get_common_fluid_params(fg_common_fluid_params* out);
int   fg_get_kernels_cache(fg_kernels_cache* out);

/* Stable in-place key/value sort of keys < 2^keyBits (1..32).  Passes are
 * chosen from keyBits; no host synchronisation. */
int fg_sort_pairs(unsigned int* keys, unsigned int* values, unsigned int n, int keyBits);
/* Same, values are taken to be 0..n-1 on input (not read). */
int fg_sort_pairs_iota(unsigned int* keys, unsigned int* values, unsigned int n, int keyBits);
/* Workspace bytes needed by fg_sort_pairs_ws for n elements. */
size_t fg_sort_workspace_size(unsigned int n, int keyBits);
int fg_sort_pairs_ws(unsigned int* keys, unsigned int* values, unsigned int n, int keyBits,
                     void* workspace, size_t workspace_bytes, int values_are_iota);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc -- fg_sort_pairs_ws without the workspace argument:
 * the same stable sort (same results), with its workspace in a file-scope
 * device array of FG_SORT_NOALLOC_SCRATCH_BYTES that every GPU the library
 * runs on holds until the process ends.  It calls no hipMalloc, hipFree or
 * allocating routine.  n is limited to FG_SORT_NOALLOC_MAX_PAIRS (larger n:
 * FG_ERR_INVALID_ARGUMENT, nothing launched).  All calls on one GPU share
 * that array and each call clears what must start at zero, so calls that
 * overlap on one GPU (concurrent streams or host threads) race. */
#define FG_SORT_NOALLOC_MAX_PAIRS 1048576u
#define FG_SORT_NOALLOC_SCRATCH_BYTES (17u * 1024u * 1024u + 64u * 1024u)
int fg_sort_pairs_noalloc(unsigned int* keys, unsigned int* values, unsigned int n, int keyBits,
                          int values_are_iota);

/* cellStart[k] := 0xFFFFFFFF for every key k in sortedKeys (the cells a
 * sorted key array occupied).  When the rest of cellStart is already
 * 0xFFFFFFFF, this resets the grid for the next cell-range pass with one
 * write per occupied cell instead of the fill of every bucket
 * (gFindCellOffsetOrderingPosVel, K_HashGrid.cu:254-255).  Keys >= numBuckets
 * are skipped. */
int fg_clear_occupied_cells(unsigned int* gridStart, const unsigned int* sortedKeys, unsigned int numParticles,
                            unsigned int numBuckets);
/* fg_calc_hash_key_only fused with fg_clear_occupied_cells: hashKey holds
 * on entry the previous step's sorted keys (whose cells are cleared in
 * gridStart) and on exit the new unsorted keys of positions. */
int fg_calc_hash_key_clear(unsigned int* hashKey, fg_float4* positions, unsigned int numParticles,
                           unsigned int numBuckets, fg_float3 cell_size, unsigned int* gridStart);
/* gFindCellOffsetOrderingPosVel with the cellStart fill optional
 * (fillStart = 0: the caller guarantees cellStart is 0xFFFFFFFF outside the
 * cells of these keys, e.g. after fg_clear_occupied_cells).  With a
 * neighbour buffer (fg_neighbor_mask_bytes, else NULL) it also writes the
 * sorted particles' FP16 cell offsets that the mixed-precision density pass
 * reads (FG_PRECISION_MIXED). */
int fg_cell_offsets_pos_vel(unsigned int* gridPosStart, unsigned int* gridPosEnd, unsigned int* hashTable,
                            unsigned int* particleID, fg_float4* particlePositions, fg_float4* orderedParticlePositions,
                            fg_float4* particleVel, fg_float4* particleOrderedVel, unsigned int numParticles,
                            unsigned int numBuckets, int fillStart, unsigned long long* neighborBuffer);

/* gCalcHashKey without the id output (the sort that follows generates the
 * identity ids itself, see fg_sort_pairs_iota). */
int fg_calc_hash_key_only(unsigned int* hashKey, fg_float4* positions, unsigned int numParticles,
                          unsigned int numBuckets, fg_float3 cell_size);

/* gSPHCalculateForces followed by gSPHIntegrate, in one pass over the sorted
 * arrays: vel[id[i]] = v_i + a_i*dt and, where pos[id[i]].w > 0,
 * pos[id[i]] = orderedPositions[i] + vel[id[i]]*(dt/scale).  Results are
 * bit-identical to the two launches in sequence. */
int fg_sph_forces_integrate(fg_float4* orderedPositions, fg_float4* orderedVelocity,
                            fg_float2* density_pressure, unsigned int* gridStart,
                            unsigned int* gridEnd, fg_float4* vel, fg_float4* pos,
                            unsigned int* particleId, unsigned int numParticles,
                            float mass, float viscosity);

/* Density and forces with per-particle neighbour bitmasks (FG_KEY_LINEAR
 * with FG_MATH_FAST; other modes ignore the masks).  The density launch
 * records, for each of a particle's 9 stencil rows, which particles of the
 * row lie within h; the force launch then evaluates only those pairs.
 * neighborMasks holds fg_neighbor_mask_bytes(n) bytes: the 9 row masks of
 * every particle, then, from the next 256-byte boundary, a 32-byte record of
 * every particle, (x, y, z, 1/rho) and (vx, vy, vz, p/rho) from the sorted
 * positions and velocities, which the force launch reads for each neighbour
 * with two 16-byte loads.  The force launch must follow the density launch on
 * the same sorted arrays.  pos may be NULL (no integration). */
size_t fg_neighbor_mask_bytes(unsigned int numParticles);
int fg_sph_density_nbr(fg_float4* orderedPositions, const fg_float4* orderedVelocity, fg_float2* mass_pressure,
                       unsigned int* gridStart, unsigned int* gridEnd, unsigned long long* neighborMasks, unsigned int numParticles,
                       float gasConstant, float restDensity, float mass);
int fg_sph_forces_integrate_nbr(fg_float4* orderedPositions, fg_float4* orderedVelocity,
                                fg_float2* density_pressure, unsigned int* gridStart, unsigned int* gridEnd,
                                const unsigned long long* neighborMasks, fg_float4* vel, fg_float4* pos,
                                unsigned int* particleId, unsigned int numParticles, float mass, float viscosity);

/* Body-sphere grid seen by the coupled fluid force pass (the static- or
 * dynamic-body arguments of gSPHCalculateForcesWithSRB, with explicit
 * counts).  A static grid has bodyVelocities NULL and numBodies 0. */
typedef struct fg_body_grid {
    const fg_float4* orderedSpheres;   /* world sphere centres in grid order, w = body index */
    const fg_float4* bodyVelocities;   /* [numBodies]; NULL for static spheres */
    const unsigned int* gridStart;     /* [numBuckets] */
    const unsigned int* gridEnd;       /* [numBuckets] */
    const fg_float4* bounds;           /* [2] expanded sphere bounds (fg_sphere_bounds) or NULL */
    fg_float3 cellSize;                /* body grid cell size (REF_XOR keys) */
    unsigned int numBuckets;
    fg_int3 relation;                  /* stencil half-width per axis */
    unsigned int numSpheres;
    unsigned int numBodies;
    float fluidRadius, sphereRadius;   /* contact at fluidRadius + sphereRadius + 0.003 */
    float kSpring, kDamping;           /* F += (kSpring*l - kDamping*dot(n, dv)) n */
} fg_body_grid;

/* Coupled fluid forces: gSPHCalculateForcesWithSRB (K_FluidSim.cu:315, :616)
 * for a static and a dynamic body grid (either may be NULL, not both; the
 * static contact is added first, as in the reference), plus the optional
 * neighbour masks of fg_sph_density_nbr, optional fused integration (pos may
 * be NULL) and sphere bounds that skip the body walk for particles that
 * cannot touch a body. */
int fg_sph_forces_coupled(fg_float4* orderedPositions, fg_float4* orderedVelocity, fg_float2* density_pressure,
                          unsigned int* gridStart, unsigned int* gridEnd, const unsigned long long* neighborMasks,
                          fg_float4* vel, fg_float4* pos, unsigned int* particleId, unsigned int numParticles,
                          float mass, float viscosity, const fg_body_grid* bodies,
                          const fg_body_grid* staticBodies);

/* gCalculateRigidBodyForcesFromFluids with explicit sphere/body/fluid counts
 * (reads are bounded by them).  The fluid grid is walked with the current
 * key mode. */
int fg_body_forces_from_fluid(fg_float4* spheres, fg_float4* bodyVelocities, fg_float4* bodyCM, fg_float4* outForces,
                              fg_float4* outTorques, unsigned int numSpheres, unsigned int numBodies, float radius,
                              fg_float4* fluidOrderedPositions, fg_float4* fluidOrderedVelocities,
                              unsigned int* fluidGridStart, unsigned int* fluidGridEnd, unsigned int numFluid,
                              float fluidRadius, fg_float3 fluidCellSize, unsigned int fluidNumBuckets,
                              fg_int3 relation, float kSpring, float kDamping);

/* gCalculateRigidBodyForcesFromRB with explicit counts: spheres (body order,
 * w = body id, any order works) against the dynamic grid `bodies` (the
 * spheres of the own body skipped; may be NULL) and the static grid
 * `staticBodies` (may be NULL, not both), contact within radius +
 * grid.sphereRadius + 0.001 over grid.relation.  bodyBounds (optional,
 * [2*numBodies], fg_body_bounds expanded by the contact distance) and the
 * static grid's bounds skip walks that cannot find a contact.  accumulate !=
 * 0 adds force and torque to outForces/outTorques (for instance the fluid's
 * part from fg_body_forces_from_fluid), else writes them. */
int fg_body_forces_from_bodies(const fg_float4* spheres, const fg_float4* bodyVelocities, const fg_float4* bodyCM,
                               fg_float4* outForces, fg_float4* outTorques, unsigned int numSpheres,
                               unsigned int numBodies, float radius, const fg_body_grid* bodies,
                               const fg_float4* bodyBounds, const fg_body_grid* staticBodies, int accumulate);

/* Per-body bounds of the sphere centres (spheres of body b occupy
 * [bodyOffsets[b], bodyOffsets[b+1])), expanded by margin: out2[2b] min,
 * out2[2b+1] max.  The broad phase of fg_body_forces_from_bodies; replaces
 * the host bounding-box test of RigidBodySimulator.cpp:334-406. */
int fg_body_bounds(const fg_float4* spheres, const unsigned int* bodyOffsets, unsigned int numBodies, float margin,
                   fg_float4* out2);

/* Per-body force and torque sums in FP64 (replaces the CUDPP segmented scan
 * of RigidBodySimulator.cpp:551-553): spheres of body b occupy
 * [bodyOffsets[b], bodyOffsets[b+1]).  accForces/accTorques (optional)
 * receive the sums rounded to float for gIntegrateRB. */
int fg_body_reduce_f64(const fg_float4* forces, const fg_float4* torques, const unsigned int* bodyOffsets,
                       unsigned int numBodies, fg_double4* outForces, fg_double4* outTorques, fg_float4* accForces,
                       fg_float4* accTorques);

/* Centre of mass (3 doubles per body) and inertia tensor about it (9 doubles,
 * row-major) in FP64 of equal point masses masses[b]/count at the sphere
 * centres of each body. */
int fg_body_mass_properties_f64(const fg_float4* worldSpheres, const unsigned int* bodyOffsets, unsigned int numBodies,
                                const double* masses, double* centreOfMass, double* inertia);

/* Bounds of all sphere centres expanded by margin: out2[0] min, out2[1] max. */
int fg_sphere_bounds(const fg_float4* spheres, unsigned int numSpheres, float margin, fg_float4* out2);

/* Inflow emission (IEmitter::update, Emitter.h:29; the reference's only
 * emitter leaves it empty, Emitter.h:46).  The particle array holds a
 * reserve of inactive slots (w <= 0) parked where no active particle
 * reaches; layers of it are activated on a lattice in the inflow window. */
typedef struct fg_inflow_desc {
    fg_float3 origin;        /* lattice site (0, 0) of the inflow window */
    fg_float3 axisU, axisV;  /* in-plane lattice axes (unit vectors) */
    fg_float3 normal;        /* inflow direction (unit vector) */
    float spacing;           /* lattice spacing, and the distance between layers (world units) */
    float speed;             /* inflow speed along normal (simulation units, as velocities) */
    unsigned int nu, nv;     /* lattice sites per layer */
    unsigned int firstSlot;  /* first reserve slot */
    unsigned int numLayers;  /* the reserve: numLayers * nu * nv slots from firstSlot */
} fg_inflow_desc;

/* Layer period in seconds: spacing * simulationScale / speed (FP64).  A
 * step moves a particle by v dt / simulation_scale world units
 * (K_FluidSim.cu:433-438), so consecutive layers are `spacing` world units
 * apart.  0 for a null descriptor or a non-positive speed. */
double fg_inflow_period(const fg_inflow_desc* desc, float simulationScale);

/* Emission at the start of a step.  With k = *stepCounter and period =
 * fg_inflow_period(desc, simulation_scale of gInitializeCommonFluidParams),
 * layers before step k = floor(k dt / period) (FP64); when that count grows
 * from k to k+1, layer l = floor(k dt / period), if l < numLayers, is
 * emitted: slot firstSlot + l nu nv + iu + nu iv gets position
 * origin + (iu spacing) axisU + (iv spacing) axisV (FP32, left to right),
 * w = 1, velocity speed * normal.  *stepCounter is then incremented.  At
 * most one layer is due per step: dt > period is refused
 * (FG_ERR_INVALID_ARGUMENT), as are missing common parameters
 * (FG_ERR_NOT_INITIALIZED).  Two launches, no host synchronisation
 * (graph-capturable). */
int fg_emit_inflow(fg_float4* positions, fg_float4* velocities, unsigned int* stepCounter, const fg_inflow_desc* desc,
                   float dt);

/* FP64 global sums of one fluid for conservation reports (no counterpart in
 * the reference).  out8 (device memory) receives, over active particles
 * (w > 0):
 *   [0] count, [1] mass, [2..4] momentum m v, [5] kinetic energy m|v|^2/2,
 *   [6] gravitational potential energy -m g.(x s), with g and the
 *       simulation scale s of the current common parameters,
 *   [7] internal energy of the equation of state p = k (rho - rho0),
 *       m k (ln(rho/rho0) + rho0/rho - 1), from the sorted densities and the
 *       active flags of the sorted positions.
 * positions/velocities in particle order; the two partial passes use a
 * caller workspace of fg_conservation_workspace_size() bytes.  The result
 * does not depend on timing (fixed reduction trees). */
size_t fg_conservation_workspace_size(void);
int fg_conservation_f64(const fg_float4* positions, const fg_float4* velocities, const fg_float4* orderedPositions,
                        const fg_float2* density_pressure, unsigned int numParticles, float mass, float restDensity,
                        float gasConstant, double* out8, void* workspace, size_t workspace_bytes);

/* Generalized winding numbers of a triangle mesh on a point lattice, for
 * sampling meshes into body sphere sets (the reference's body particle
 * files, Simulators/RigidBody.cpp:55-108, name the mesh they came from).
 * Point (ix, iy, iz), 0 <= i < n per axis, is origin + i * spacing per axis
 * (product and sum each rounded to FP32); winding[ix + nx (iy + ny iz)]
 * receives sum_t Omega_t / (4 pi), Omega_t the signed solid angle of
 * triangle t (corners vertices[triangles[3t..3t+2]], counter-clockwise seen
 * from outside), each term 2 atan2(a.(b x c), |a||b||c| + (a.b)|c| +
 * (b.c)|a| + (c.a)|b|) in FP32 with a, b, c the corners minus the point,
 * summed in FP64.  A triangle with a corner index >= numVertices adds 0.
 * Device pointers; at most 2^34 points. */
int fg_mesh_winding_lattice(const fg_float3* vertices, unsigned int numVertices, const unsigned int* triangles,
                            unsigned int numTriangles, fg_float3 origin, float spacing, unsigned int nx,
                            unsigned int ny, unsigned int nz, float* winding);

/* ------------------------------------------------------------------------ */
/* 3. Batched scenes (no counterpart in the reference)                       */
/* ------------------------------------------------------------------------ */
/* Many independent fluids advance with one launch per stage.  Particles are
 * stored scene-major: scene s owns [particleBegin, particleBegin +
 * particleCount) of every particle array.  Each scene has its own LINEAR
 * grid (common.bucket_size, cell_size, volume_min); its buckets occupy
 * [keyBase, keyBase + bucket count) of one shared key space, with keyBase
 * increasing in scene order, so one stable sort of the shared keys sorts
 * every scene inside its own particle range and the grid arrays
 * (gridStart/gridEnd of the total bucket count) serve all scenes.
 * sceneOf[i] is the scene of particle i (the same for sorted and unsorted
 * order, since the sort keeps scene ranges); particles whose sceneOf is not
 * below numScenes are left untouched.  Per-scene arithmetic is that of
 * the single-fluid FG_KEY_LINEAR / FG_MATH_FAST path. */
typedef struct fg_scene_desc {
    fg_common_fluid_params common;     /* grid, box, gravity, dt, scale, h of this scene */
    fg_kernels_cache kernels;
    float mass, gasConstant, restDensity, viscosity;
    unsigned int keyBase;              /* first bucket of this scene in the shared key space */
    unsigned int particleBegin;
    unsigned int particleCount;
    unsigned int reserved[4];
} fg_scene_desc;                       /* 160 bytes */

/* key[i] = keyBase + LINEAR key of positions[i] in its scene's grid. */
int fg_batch_hash(const fg_scene_desc* scenes, unsigned int numScenes, const unsigned int* sceneOf,
                  const fg_float4* positions, unsigned int* hashKey, unsigned int numParticles);
/* fg_sph_density_nbr per scene (neighbour masks always written).  `scenes`
 * is device memory, so the launcher cannot check it on the host: a scene
 * whose cells are smaller than its kernel radius on any axis (the 27-cell
 * stencil would miss neighbours) gets NaN (rho, p) for its particles, and the
 * force pass writes NaN velocities for them, instead of short sums.
 * fg_bsim_create refuses such scenes before anything runs. */
int fg_batch_density_nbr(const fg_scene_desc* scenes, unsigned int numScenes, const unsigned int* sceneOf,
                         fg_float4* orderedPositions, const fg_float4* orderedVelocity, fg_float2* mass_pressure,
                         unsigned int* gridStart, unsigned int* gridEnd, unsigned long long* neighborMasks,
                         unsigned int numParticles);
/* fg_sph_forces_integrate_nbr per scene (pos may be NULL: no integration). */
int fg_batch_forces_integrate_nbr(const fg_scene_desc* scenes, unsigned int numScenes,
                                  const unsigned int* sceneOf, fg_float4* orderedPositions, fg_float4* orderedVelocity,
                                  fg_float2* density_pressure, unsigned int* gridStart, unsigned int* gridEnd,
                                  const unsigned long long* neighborMasks, fg_float4* vel, fg_float4* pos,
                                  unsigned int* particleId, unsigned int numParticles);

/* Device memory and timing helpers (thin wrappers over the HIP runtime). */
int fg_device_count(int* count);
int fg_device_name(int device, char* buf, size_t len);
int fg_malloc(void** ptr, size_t bytes);
int fg_free(void* ptr);
int fg_memcpy_htod(void* dst, const void* src, size_t bytes);
int fg_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int fg_memcpy_dtod(void* dst, const void* src, size_t bytes);
int fg_memset_u32(unsigned int* dst, unsigned int value, size_t count); /* fill kernel */
int fg_copy_bytes(void* dst, const void* src, size_t bytes);          /* streaming copy kernel, 16-B aligned */
int fg_synchronize(void);                  /* the selected stream */
int fg_event_create(void** ev);
int fg_event_destroy(void* ev);
int fg_event_record(void* ev);             /* on the selected stream */
int fg_event_record_stream(void* ev, void* stream);  /* on a given hipStream_t (e.g. fg_sim_stream()) */
int fg_event_elapsed_ms(void* start, void* stop, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* FLUIDGRID_AMD_H */
