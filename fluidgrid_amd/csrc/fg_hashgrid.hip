// fg_hashgrid.hip -- neighbourhood grid: cell keys, body-sphere transform,
// cell start/end and the SoA reorder (the work of GPU_Kernels/K_HashGrid.cu).
//
// All kernels are streaming: one lane per particle, 256-lane workgroups
// (four wave64s), 16-byte loads and stores for the float4 arrays.
#include "fg_internal.hpp"
#include "fg_sph_mixed.hpp"
#include "fg_quat.hpp"

namespace fg {

GridDesc fluid_grid_desc(const fg_common_fluid_params& cp, fg_float3 cs, uint32_t nb, int mode) {
    GridDesc g;
    g.vmin = make_float3(cp.volume_min.x, cp.volume_min.y, cp.volume_min.z);
    g.cs = make_float3(cs.x, cs.y, cs.z);
    g.dims = make_int3(cp.bucket_size.x, cp.bucket_size.y, cp.bucket_size.z);
    g.nb = nb;
    g.mode = mode;
    g.key_base = 0;
    return g;
}

bool validate_grid(const GridDesc& g, const char* who) {
    if (g.nb == 0) {
        set_error(FG_ERR_INVALID_ARGUMENT, "%s: numBuckets is 0", who);
        return false;
    }
    if (g.mode == FG_KEY_LINEAR) {
        uint64_t prod = (uint64_t)g.dims.x * (uint64_t)g.dims.y * (uint64_t)g.dims.z;
        if (g.dims.x <= 0 || g.dims.y <= 0 || g.dims.z <= 0 || prod != g.nb) {
            set_error(FG_ERR_INVALID_ARGUMENT,
                      "%s: LINEAR keys need numBuckets == bucket_size.x*y*z (%u vs %llu); call "
                      "gInitializeCommonFluidParams first",
                      who, g.nb, (unsigned long long)prod);
            return false;
        }
    }
    return true;
}

// K_HashGrid.cu:26-38
__global__ void __launch_bounds__(256) k_hash(uint32_t* __restrict__ key, uint32_t* __restrict__ id,
                                              const float4* __restrict__ pos, uint32_t n, GridDesc g) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    key[i] = grid_key(p.x, p.y, p.z, g);
    if (id) id[i] = i;
}

// k_hash without ids, where key[] holds on entry the previous sorted keys:
// before its key is replaced, each thread clears cellStart of its old key
// when that key starts a run inside the block (runs that cross a block
// boundary are cleared once per block, which is harmless).  Neighbouring old
// keys are exchanged through LDS, so no thread reads a key another thread
// overwrites.
__global__ void __launch_bounds__(256) k_hash_clear(uint32_t* __restrict__ key, const float4* __restrict__ pos,
                                                    uint32_t n, GridDesc g, uint32_t* __restrict__ start,
                                                    uint32_t nb) {
    __shared__ uint32_t s_old[256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t old = i < n ? key[i] : kEmpty;
    s_old[threadIdx.x] = old;
    __syncthreads();
    if (i >= n) return;
    if (old < nb && (threadIdx.x == 0 || s_old[threadIdx.x - 1] != old)) start[old] = kEmpty;
    const float4 p = pos[i];
    key[i] = grid_key(p.x, p.y, p.z, g);
}

// K_HashGrid.cu:43-63 with K_MathUtils.cu:160-181
__global__ void __launch_bounds__(256) k_hash_with_offset(uint32_t* __restrict__ key, uint32_t* __restrict__ id,
                                                          const float4* __restrict__ offset,
                                                          const float4* __restrict__ orient,
                                                          const float4* __restrict__ rel, float4* __restrict__ out,
                                                          uint32_t n, GridDesc g) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = rel[i];
    const int b = f2i_sat(p.w);
    const float4 a = world_sphere(p, orient[b], offset[b]);
    id[i] = i;
    out[i] = a;
    key[i] = grid_key(a.x, a.y, a.z, g);
}

// K_HashGrid.cu:70-133 / 140-197.  cellStart[k] = first sorted index of key
// k, cellEnd[k] = one past the last; sorted arrays gathered through id.
template <bool WITH_VEL>
__global__ void __launch_bounds__(256) k_cell_offsets(uint32_t* __restrict__ start, uint32_t* __restrict__ end,
                                                      const uint32_t* __restrict__ key,
                                                      const uint32_t* __restrict__ id,
                                                      const float4* __restrict__ pos, float4* __restrict__ spos,
                                                      const float4* __restrict__ vel, float4* __restrict__ svel,
                                                      uint32_t n, uint32_t nb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // Keys and ids outside their ranges (inputs not produced by the hash
    // and sort) are skipped rather than written out of bounds.
    const uint32_t h = key[i];
    if (i == 0) {
        if (h < nb) start[h] = 0;
    } else {
        const uint32_t prev = key[i - 1];
        if (h != prev) {
            if (h < nb) start[h] = i;
            if (prev < nb) end[prev] = i;
        }
    }
    if (i == n - 1 && h < nb) end[h] = n;
    const uint32_t src = id[i];
    if (src >= n) return;
    spos[i] = pos[src];
    if (WITH_VEL) svel[i] = vel[src];
}

// k_cell_offsets<true> that also writes the FP16 pair records of the mixed
// precision form (fg_sph_mixed.hpp): record i holds the offsets of sorted
// particles i and i+1.  Particle i+1's offsets come from the next lane; the
// last lane of a wavefront gathers that particle itself.
__global__ void __launch_bounds__(256) k_cell_offsets_half(uint32_t* __restrict__ start, uint32_t* __restrict__ end,
                                                           const uint32_t* __restrict__ key,
                                                           const uint32_t* __restrict__ id,
                                                           const float4* __restrict__ pos, float4* __restrict__ spos,
                                                           const float4* __restrict__ vel, float4* __restrict__ svel,
                                                           uint32_t n, uint32_t nb, uint4* __restrict__ hrec,
                                                           fg_common_fluid_params cp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const HalfGeom g = half_geom(cp);
    float3 o = make_float3(0.0f, 0.0f, 0.0f);
    if (valid) {
        const uint32_t h = key[i];
        if (i == 0) {
            if (h < nb) start[h] = 0;
        } else {
            const uint32_t prev = key[i - 1];
            if (h != prev) {
                if (h < nb) start[h] = i;
                if (prev < nb) end[prev] = i;
            }
        }
        if (i == n - 1 && h < nb) end[h] = n;
        const uint32_t src = id[i];
        if (src < n) {
            const float4 p = pos[src];
            spos[i] = p;
            svel[i] = vel[src];
            o = half_offsets(p, cp, g);
        }
    }
    const uint32_t xy = pack_half2(o.x, o.y), zz = pack_half2(o.z, 0.0f);
    uint32_t nxy = (uint32_t)__shfl_down((int)xy, 1), nzz = (uint32_t)__shfl_down((int)zz, 1);
    if (__lane_id() == 63u) {
        float3 q = make_float3(0.0f, 0.0f, 0.0f);
        if (i + 1 < n) {
            const uint32_t src = id[i + 1];
            if (src < n) q = half_offsets(pos[src], cp, g);
        }
        nxy = pack_half2(q.x, q.y);
        nzz = pack_half2(q.z, 0.0f);
    } else if (i + 1 >= n) {
        nxy = nzz = 0u;
    }
    if (valid)
        hrec[i] = make_uint4(__builtin_amdgcn_perm(nxy, xy, 0x05040100u), __builtin_amdgcn_perm(nxy, xy, 0x07060302u),
                             __builtin_amdgcn_perm(nzz, zz, 0x05040100u), 0u);
    if (i == 0) {
        hrec[n] = make_uint4(0u, 0u, 0u, 0u);
        hrec[n + 1] = make_uint4(0u, 0u, 0u, 0u);
    }
}

// cellStart[k] = kEmpty for every key k of a sorted key array (the cells
// that array occupied): with the rest of cellStart already kEmpty, this
// clears it for the next cell-range pass at a cost of one write per
// occupied cell instead of one per bucket.
__global__ void __launch_bounds__(256) k_clear_occupied(uint32_t* __restrict__ start,
                                                        const uint32_t* __restrict__ key, uint32_t n, uint32_t nb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = key[i];
    if (k < nb && (i == 0 || key[i - 1] != k)) start[k] = kEmpty;
}

}  // namespace fg

using namespace fg;

extern "C" {

// K_HashGrid.cu:207-216
void gCalcHashKey(unsigned int* hashKey, unsigned int* particleID, fg_float4* positions, unsigned int numParticles,
                  unsigned int numBuckets, fg_float3 cell_size) {
    if (numParticles == 0) return;
    if (!hashKey || !particleID || !positions) {
        set_error(FG_ERR_INVALID_ARGUMENT, "gCalcHashKey: null pointer");
        return;
    }
    if (!(cell_size.x > 0.0f && cell_size.y > 0.0f && cell_size.z > 0.0f)) {
        set_error(FG_ERR_INVALID_ARGUMENT,
                  "gCalcHashKey: cell_size (%.9g, %.9g, %.9g) must be positive (numParticles=%u numBuckets=%u)",
                  (double)cell_size.x, (double)cell_size.y, (double)cell_size.z, numParticles, numBuckets);
        return;
    }
    const State& st = state();
    GridDesc g = fluid_grid_desc(st.common, cell_size, numBuckets, st.key_mode);
    if (!validate_grid(g, "gCalcHashKey")) return;
    hipLaunchKernelGGL(k_hash, dim3(grid_for(numParticles, 256)), dim3(256), 0, st.stream, hashKey, particleID,
                       FG_F4(positions), numParticles, g);
    check_launch("gCalcHashKey");
}

// Keys only (ids are implied 0..n-1 by the sort that follows).
int fg_calc_hash_key_only(unsigned int* hashKey, fg_float4* positions, unsigned int numParticles,
                          unsigned int numBuckets, fg_float3 cell_size) {
    if (numParticles == 0) return FG_OK;
    const State& st = state();
    GridDesc g = fluid_grid_desc(st.common, cell_size, numBuckets, st.key_mode);
    if (!validate_grid(g, "fg_calc_hash_key_only")) return FG_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(k_hash, dim3(grid_for(numParticles, 256)), dim3(256), 0, st.stream, hashKey,
                       (uint32_t*)nullptr, FG_F4(positions), numParticles, g);
    return check_launch("fg_calc_hash_key_only") ? FG_OK : fg_peek_error();
}

int fg_calc_hash_key_clear(unsigned int* hashKey, fg_float4* positions, unsigned int numParticles,
                           unsigned int numBuckets, fg_float3 cell_size, unsigned int* gridStart) {
    if (numParticles == 0) return FG_OK;
    if (!hashKey || !positions || !gridStart) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_calc_hash_key_clear: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    const State& st = state();
    GridDesc g = fluid_grid_desc(st.common, cell_size, numBuckets, st.key_mode);
    if (!validate_grid(g, "fg_calc_hash_key_clear")) return FG_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(k_hash_clear, dim3(grid_for(numParticles, 256)), dim3(256), 0, st.stream, hashKey,
                       FG_F4(positions), numParticles, g, gridStart, numBuckets);
    return check_launch("fg_calc_hash_key_clear") ? FG_OK : fg_peek_error();
}

// K_HashGrid.cu:221-233
void gCalcHashKeyWithOffset(unsigned int* _out_hashKey, unsigned int* _out_particleID, fg_float4* _in_vectorOffset,
                            fg_float4* _in_orientation, fg_float4* _in_vectorPos, fg_float4* _out_positions,
                            unsigned int _in_numParticles, fg_float3 _in_cellSize, int _in_hashTableSize) {
    if (_in_numParticles == 0) return;
    if (!_out_hashKey || !_out_particleID || !_in_vectorOffset || !_in_orientation || !_in_vectorPos ||
        !_out_positions || _in_hashTableSize <= 0) {
        set_error(FG_ERR_INVALID_ARGUMENT, "gCalcHashKeyWithOffset: invalid argument");
        return;
    }
    const State& st = state();
    // Body grids are keyed with the reference hash whatever the fluid mode:
    // their bucket count is independent of the fluid grid dimensions.
    GridDesc g = fluid_grid_desc(st.common, _in_cellSize, (uint32_t)_in_hashTableSize, FG_KEY_REF_XOR);
    hipLaunchKernelGGL(k_hash_with_offset, dim3(grid_for(_in_numParticles, 256)), dim3(256), 0, st.stream,
                       _out_hashKey, _out_particleID, FG_F4(_in_vectorOffset), FG_F4(_in_orientation),
                       FG_F4(_in_vectorPos), FG_F4(_out_positions), _in_numParticles, g);
    check_launch("gCalcHashKeyWithOffset");
}

// K_HashGrid.cu:238-281
void gFindCellOffsetOrderingPosVel(unsigned int* gridPosStart, unsigned int* gridPosEnd, unsigned int* hashTable,
                                   unsigned int* particleID, fg_float4* particlePositions,
                                   fg_float4* orderedParticlePositions, fg_float4* particleVel,
                                   fg_float4* particleOrderedVel, unsigned int numParticles,
                                   unsigned int numBuckets) {
    if (!gridPosStart || (numParticles && (!gridPosEnd || !hashTable || !particleID || !particlePositions ||
                                           !orderedParticlePositions || !particleVel || !particleOrderedVel))) {
        set_error(FG_ERR_INVALID_ARGUMENT, "gFindCellOffsetOrderingPosVel: null pointer");
        return;
    }
    const State& st = state();
    launch_fill_u32(gridPosStart, kEmpty, numBuckets, st.stream);
    if (numParticles)
        hipLaunchKernelGGL(k_cell_offsets<true>, dim3(grid_for(numParticles, 256)), dim3(256), 0, st.stream,
                           gridPosStart, gridPosEnd, hashTable, particleID, FG_F4(particlePositions),
                           FG_F4(orderedParticlePositions), FG_F4(particleVel), FG_F4(particleOrderedVel),
                           numParticles, numBuckets);
    check_launch("gFindCellOffsetOrderingPosVel");
}

int fg_clear_occupied_cells(unsigned int* gridStart, const unsigned int* sortedKeys, unsigned int numParticles,
                            unsigned int numBuckets) {
    if (numParticles == 0) return FG_OK;
    if (!gridStart || !sortedKeys) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_clear_occupied_cells: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    hipLaunchKernelGGL(k_clear_occupied, dim3(grid_for(numParticles, 256)), dim3(256), 0, state().stream, gridStart,
                       sortedKeys, numParticles, numBuckets);
    return check_launch("fg_clear_occupied_cells") ? FG_OK : fg_peek_error();
}

int fg_cell_offsets_pos_vel(unsigned int* gridPosStart, unsigned int* gridPosEnd, unsigned int* hashTable,
                            unsigned int* particleID, fg_float4* particlePositions, fg_float4* orderedParticlePositions,
                            fg_float4* particleVel, fg_float4* particleOrderedVel, unsigned int numParticles,
                            unsigned int numBuckets, int fillStart, unsigned long long* neighborBuffer) {
    if (!gridPosStart || (numParticles && (!gridPosEnd || !hashTable || !particleID || !particlePositions ||
                                           !orderedParticlePositions || !particleVel || !particleOrderedVel))) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_cell_offsets_pos_vel: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    const State& st = state();
    if (neighborBuffer && !st.have_common) {
        set_error(FG_ERR_NOT_INITIALIZED, "fg_cell_offsets_pos_vel: FP16 offsets need gInitializeCommonFluidParams");
        return FG_ERR_NOT_INITIALIZED;
    }
    if (fillStart) launch_fill_u32(gridPosStart, kEmpty, numBuckets, st.stream);
    if (numParticles && neighborBuffer) {
        uint4* hrec = reinterpret_cast<uint4*>(reinterpret_cast<char*>(neighborBuffer) + nbr_half_offset(numParticles));
        hipLaunchKernelGGL(k_cell_offsets_half, dim3(grid_for(numParticles, 256)), dim3(256), 0, st.stream,
                           gridPosStart, gridPosEnd, hashTable, particleID, FG_F4(particlePositions),
                           FG_F4(orderedParticlePositions), FG_F4(particleVel), FG_F4(particleOrderedVel),
                           numParticles, numBuckets, hrec, st.common);
        note_half_records(neighborBuffer, numParticles);
    } else if (numParticles) {
        hipLaunchKernelGGL(k_cell_offsets<true>, dim3(grid_for(numParticles, 256)), dim3(256), 0, st.stream,
                           gridPosStart, gridPosEnd, hashTable, particleID, FG_F4(particlePositions),
                           FG_F4(orderedParticlePositions), FG_F4(particleVel), FG_F4(particleOrderedVel),
                           numParticles, numBuckets);
    }
    return check_launch("fg_cell_offsets_pos_vel") ? FG_OK : fg_peek_error();
}

// K_HashGrid.cu:287-318
void gFindCellOffsetOrderingPos(unsigned int* gridPosStart, unsigned int* gridPosEnd, unsigned int* hashTable,
                                unsigned int* particleID, fg_float4* particlePositions,
                                fg_float4* orderedParticlePositions, unsigned int numParticles,
                                unsigned int numBuckets) {
    if (!gridPosStart || (numParticles && (!gridPosEnd || !hashTable || !particleID || !particlePositions ||
                                           !orderedParticlePositions))) {
        set_error(FG_ERR_INVALID_ARGUMENT, "gFindCellOffsetOrderingPos: null pointer");
        return;
    }
    const State& st = state();
    launch_fill_u32(gridPosStart, kEmpty, numBuckets, st.stream);
    if (numParticles)
        hipLaunchKernelGGL(k_cell_offsets<false>, dim3(grid_for(numParticles, 256)), dim3(256), 0, st.stream,
                           gridPosStart, gridPosEnd, hashTable, particleID, FG_F4(particlePositions),
                           FG_F4(orderedParticlePositions), (const float4*)nullptr, (float4*)nullptr,
                           numParticles, numBuckets);
    check_launch("gFindCellOffsetOrderingPos");
}

}  // extern "C"
