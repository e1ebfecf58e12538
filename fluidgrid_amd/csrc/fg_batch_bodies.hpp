// fg_batch_bodies.hpp -- launchers of the rigid-body stages of the batched
// step (library-internal: called by host/BatchedSimulator.cpp).  Every
// per-element computation is the single-fluid one (fg_rigid.hip,
// fg_sph_fast.hpp) with the parameters of the element's scene, so each
// scene's bodies and fluid advance exactly as in a single-fluid simulator.
#pragma once
#include "fg_internal.hpp"

namespace fg {

// World positions and body-grid keys of all body spheres (gCalcHashKeyWithOffset per scene).
int batch_body_transform(uint32_t* key, uint32_t* id, const float4* cm, const float4* orient, const float4* rel,
                         float4* world, uint32_t m, const uint32_t* body_scene, const BatchBodyScene* scenes);
// Per scene: bounds of its sphere centres grown by bounds_margin (fg_sphere_bounds), into out[2s], out[2s+1].
int batch_scene_bounds(const float4* world, const BatchBodyScene* scenes, uint32_t nscenes, float4* out);
// Per body: bounds grown by its scene's pair_margin (fg_body_bounds).
int batch_body_bounds(const float4* world, const uint32_t* offsets, const uint32_t* body_scene,
                      const BatchBodyScene* scenes, uint32_t nbodies, float4* out);
// Per body sphere: penalty force and torque from its scene's fluid (fg_body_forces_from_fluid), then, in
// scenes with two or more bodies, from the other bodies of the scene added on (fg_body_forces_from_bodies).
int batch_body_forces(const float4* world, const float4* vel, const float4* cm, float4* out_f, float4* out_t,
                      uint32_t m, uint32_t nbodies, const uint32_t* body_scene, const BatchBodyScene* scenes,
                      const float4* fluid_pos, const float4* fluid_vel, const uint32_t* fluid_start,
                      const uint32_t* fluid_end, uint32_t nfluid, const float4* body_bounds);
// Per body: IntegrateRBK with its scene's dt, gravity, box and velocity limit (gIntegrateRB).
int batch_integrate_bodies(float4* cm, const float* mass, float4* vel, const float4* force, uint32_t nbodies,
                           const uint32_t* body_scene, const BatchBodyScene* scenes, float4* L, const float4* torque,
                           const float4* inv_inertia, float4* orient);
// Persistent form of the density and force launches: a grid of as many
// workgroups as the GPU holds at once takes work items (per_item blocks'
// worth of consecutive sorted particles) from counters instead of one
// workgroup per block.  The items are split into kQueueSlots slices, one per
// XCD (workgroups are dispatched to the XCDs round-robin), each with its own
// counter; a workgroup whose slice is exhausted takes from the next slices.
// counters: kQueueSlots * kQueueStride words owned by the caller (cleared by
// each launch); null selects the plain launch.  Results are the same bits.
constexpr uint32_t kQueueSlots = 8;
constexpr uint32_t kQueueStride = 16;  // one 64-byte line per counter
struct WorkQueue {
    uint32_t* counters;
    uint32_t per_item;
};

// fg_batch_density_nbr / fg_batch_forces_integrate_nbr, optionally persistent.
int batch_density(const fg_scene_desc* scenes, uint32_t nscenes, const uint32_t* scene_of, float4* sp,
                  const float4* sv, float2* rho_p, uint32_t* start, uint32_t* end, unsigned long long* masks,
                  uint32_t n, WorkQueue q);
int batch_forces(const fg_scene_desc* scenes, uint32_t nscenes, const uint32_t* scene_of, float4* sp, float4* sv,
                 float2* rho_p, uint32_t* start, uint32_t* end, const unsigned long long* masks, float4* vel,
                 float4* pos, uint32_t* id, uint32_t n, WorkQueue q);
// fg_batch_forces_integrate_nbr with each scene's bodies (fg_sph_forces_coupled per scene; scenes
// without bodies take the uncoupled expression).
int batch_forces_coupled(const fg_scene_desc* scenes, uint32_t nscenes, const uint32_t* scene_of, float4* sp,
                         float4* sv, float2* rho_p, uint32_t* start, uint32_t* end, const unsigned long long* masks,
                         float4* vel, float4* pos, uint32_t* id, uint32_t n, const BatchBodyScene* bodies,
                         WorkQueue q = WorkQueue{nullptr, 1});
// The same in two passes with the same bits: the force pass (compiled
// without the contact walk) finishes the particles outside every body
// grid's bounds and defers the others (dc: n entries each, count cleared
// here), which a second pass finishes with the contact walk.
int batch_forces_deferred(const fg_scene_desc* scenes, uint32_t nscenes, const uint32_t* scene_of, float4* sp,
                          float4* sv, float2* rho_p, uint32_t* start, uint32_t* end,
                          const unsigned long long* masks, float4* vel, float4* pos, uint32_t* id, uint32_t n,
                          const BatchBodyScene* bodies, DeferredContact dc, WorkQueue q = WorkQueue{nullptr, 1});

}  // namespace fg
