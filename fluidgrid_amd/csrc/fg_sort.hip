// fg_sort.hip -- stable LSD radix sort of (u32 key, u32 value) pairs.
//
// Replaces the reference's thrust::sort_by_key (GPU_Kernels/K_FluidSim.cu:685-691)
// and the CUDPP key/value radix sorts (Simulators/RigidBodySimulator.cpp:195,
// 469, 542).  Ties keep their input order, so the result is the unique stable
// permutation of the keys.
//
// Design (wave64, gfx950): a 256-lane workgroup owns a tile of 2048 or 4096
// consecutive elements (8 or 16 per lane, by size); each digit pass is three
// launches,
//  * upsweep: per-tile digit counts (per-wave LDS histograms; a run of equal
//    digits in a wavefront is counted once, by its first lane), stored
//    digit-major;
//  * scan: one workgroup per digit turns its row of tile counts into
//    exclusive offsets within the digit and records the digit's total;
//  * downsweep: the tile is ranked stably with wave-wide ballot matching
//    (peers with an equal digit found bit by bit), its global offsets are
//    the digit base (an exclusive scan of the totals) plus the tile's offset
//    within the digit, and it is staged in LDS in digit order and written
//    out as coalesced runs.
// No workgroup waits for another (a single-sweep pass with decoupled
// look-back chains tiles: at 1M keys the chain, not the bytes, set its
// time).  Digit widths are chosen from the key width (<= 8 bits per pass),
// so 16-bit keys take two passes, 22-bit keys three, 25/27-bit keys four.
#include <mutex>

#include "fg_internal.hpp"

namespace fg {

constexpr int kSortThreads = 256;
// Items per lane: 8 (2048-element tiles) up to kSmallTileMaxN keys, where
// twice as many tiles fill the GPU better and each tile's chain of loads,
// ranking and stores is shorter; 16 (4096-element tiles) above, where the
// per-tile counts and their scan would grow.  Measured per sort (22-bit keys
// at 1M: 63 us against 75; 25-bit at 8M: 230 against 227; 27-bit at 33.5M:
// 1023 against 849).
constexpr int kSortItemsSmall = 8;
constexpr int kSortItemsLarge = 16;
constexpr uint32_t kSmallTileMaxN = 1u << 22;
static int sort_items(uint32_t n) { return n <= kSmallTileMaxN ? kSortItemsSmall : kSortItemsLarge; }
static uint32_t sort_tiles(uint32_t n) {
    const uint32_t tile = (uint32_t)(kSortThreads * sort_items(n));
    return (n + tile - 1) / tile;
}
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kMaxPasses = 4;
constexpr int kMaxRadix = 256;

struct SortPlan {
    int passes;
    int shift[kMaxPasses];
    int bits[kMaxPasses];
};

constexpr int kMinDigitBits = 6;

// Passes for keyBits-bit keys with digits of at most digitBits bits (6..8),
// never more than kMaxPasses; bits are spread evenly over the passes.
static int plan_passes(int keyBits, int digitBits) {
    if (digitBits < kMinDigitBits) digitBits = kMinDigitBits;
    if (digitBits > 8) digitBits = 8;
    int p = (keyBits + digitBits - 1) / digitBits;
    const int pmin = (keyBits + 7) / 8;
    if (p > kMaxPasses) p = kMaxPasses;
    if (p < pmin) p = pmin;
    return p;
}

static SortPlan make_plan(int keyBits, int digitBits = 8) {
    SortPlan p{};
    if (keyBits < 1) keyBits = 1;
    if (keyBits > 32) keyBits = 32;
    p.passes = plan_passes(keyBits, digitBits);
    const int base = keyBits / p.passes, rem = keyBits % p.passes;
    int s = 0;
    for (int i = 0; i < p.passes; i++) {
        p.bits[i] = base + (i < rem ? 1 : 0);
        p.shift[i] = s;
        s += p.bits[i];
    }
    return p;
}

struct SortLayout {
    size_t counters, totals, counts, t1k, t1v, t2k, t2v, total;
};

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Workspace layout for n keys of keyBits bits, large enough for every digit
// width the tuner may select: a few counters, the digit totals and the
// tile counts of one pass (passes run one after another), and the ping-pong
// key/value buffers.
static SortLayout make_layout(uint32_t n, int keyBits) {
    SortLayout L{};
    if (keyBits < 1) keyBits = 1;
    if (keyBits > 32) keyBits = 32;
    const int max_passes = plan_passes(keyBits, kMinDigitBits);
    const size_t tiles = sort_tiles(n);
    size_t o = 0;
    L.counters = o; o = align256(o + sizeof(uint32_t) * (kMaxPasses + 1));
    L.totals = o; o = align256(o + sizeof(uint32_t) * kMaxRadix);
    L.counts = o; o = align256(o + sizeof(uint32_t) * tiles * kMaxRadix);
    L.t1k = o; o = align256(o + sizeof(uint32_t) * (size_t)n);
    L.t1v = o; o = align256(o + sizeof(uint32_t) * (size_t)n);
    if (max_passes >= 3) {
        L.t2k = o; o = align256(o + sizeof(uint32_t) * (size_t)n);
        L.t2v = o; o = align256(o + sizeof(uint32_t) * (size_t)n);
    } else {
        L.t2k = L.t2v = 0;
    }
    L.total = o;
    return L;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const uint32_t lane = __lane_id();
    return (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
}

// Upsweep: digit counts of tile blockIdx.x (ITEMS keys per lane), into
// counts[d * tiles + tile].
template <int ITEMS>
__global__ void __launch_bounds__(kSortThreads) k_sort_upsweep(const uint32_t* __restrict__ keys, uint32_t n,
                                                               int shift, int bits, uint32_t* __restrict__ counts,
                                                               uint32_t tiles) {
    __shared__ uint32_t s_cnt[kSortWaves][kMaxRadix];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < kSortWaves * kMaxRadix; i += kSortThreads) (&s_cnt[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = blockIdx.x;
    const uint32_t dmask = (1u << bits) - 1u;
    const uint32_t wbase = tile * (uint32_t)(kSortThreads * ITEMS) + (uint32_t)wave * 64u * ITEMS;
    uint32_t key[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t idx = wbase + (uint32_t)k * 64u + (uint32_t)lane;
        key[k] = idx < n ? keys[idx] : 0u;
    }
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t idx = wbase + (uint32_t)k * 64u + (uint32_t)lane;
        const bool valid = idx < n;
        const uint32_t d = valid ? (key[k] >> shift) & dmask : 0xFFFFFFFFu;
        const uint32_t prev = (uint32_t)__shfl_up((int)d, 1);
        const bool head = valid && (lane == 0 || prev != d);
        const uint64_t heads = __ballot(head) | __ballot(!valid);  // an invalid lane ends the run before it
        if (head) {
            const uint64_t after = lane == 63 ? 0ull : (heads >> (lane + 1));
            const uint32_t len = after ? (uint32_t)__ffsll((long long)after) : (uint32_t)(64 - lane);
            atomicAdd(&s_cnt[wave][d], len);
        }
    }
    __syncthreads();
    const uint32_t radix = 1u << bits;
    for (uint32_t d = (uint32_t)tid; d < radix; d += kSortThreads) {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; w++) c += s_cnt[w][d];
        counts[(size_t)d * tiles + tile] = c;
    }
}

// Exclusive scan of v over the workgroup's 256 lanes (in lane order);
// *total receives the sum.  Uses s_w[kSortWaves].
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_w, uint32_t* total) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t before = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; w++) {
        const uint32_t t = s_w[w];
        before += w < wave ? t : 0u;
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return before + x - v;
}

// Scan: row blockIdx.x (one digit) of the tile counts becomes the exclusive
// offsets of the tiles within that digit; totals[d] its sum.  Lane t owns
// the run of `per` consecutive tiles starting at t*per: one workgroup scan
// of the run sums, then each lane writes its run's offsets.
__global__ void __launch_bounds__(kSortThreads) k_sort_scan(uint32_t* __restrict__ counts, uint32_t tiles,
                                                            uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_w[kSortWaves];
    uint32_t* row = counts + (size_t)blockIdx.x * tiles;
    const uint32_t per = (tiles + kSortThreads - 1) / kSortThreads;
    const uint32_t t0 = threadIdx.x * per;
    const uint32_t t1 = min(t0 + per, tiles);
    uint32_t sum = 0;
#pragma unroll 4
    for (uint32_t t = t0; t < t1; t++) sum += row[t];
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, s_w, &total);
    for (uint32_t t = t0; t < t1; t++) {
        const uint32_t c = row[t];
        row[t] = run;
        run += c;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

// Downsweep of one digit pass.  Element e of tile T is base + w*64*ITEMS +
// k*64 + lane, so (wave, item, lane) order is index order and the ranking
// is stable.
template <int ITEMS>
__global__ void __launch_bounds__(kSortThreads) k_sort_pass(const uint32_t* __restrict__ kin,
                                                            const uint32_t* __restrict__ vin,
                                                            uint32_t* __restrict__ kout,
                                                            uint32_t* __restrict__ vout, uint32_t n, int shift,
                                                            int bits, const uint32_t* __restrict__ counts,
                                                            uint32_t tiles, const uint32_t* __restrict__ totals,
                                                            int values_iota) {
    constexpr int TILE = kSortThreads * ITEMS;
    __shared__ uint32_t s_keys[TILE];
    __shared__ uint32_t s_vals[TILE];
    __shared__ uint32_t s_wcnt[kSortWaves][kMaxRadix];
    __shared__ uint32_t s_tstart[kMaxRadix];
    __shared__ uint32_t s_gbase[kMaxRadix];
    __shared__ uint32_t s_w[kSortWaves];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t radix = 1u << bits, dmask = radix - 1u;
    for (int i = tid; i < kSortWaves * kMaxRadix; i += kSortThreads) (&s_wcnt[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * (uint32_t)TILE;
    const uint32_t wbase = base + (uint32_t)wave * 64u * ITEMS;

    uint32_t key[ITEMS], val[ITEMS], rnk[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t idx = wbase + (uint32_t)k * 64u + (uint32_t)lane;
        const bool valid = idx < n;
        key[k] = valid ? kin[idx] : 0u;
        val[k] = valid ? (values_iota ? idx : vin[idx]) : 0u;
    }
    // this tile's offset within each digit and the digit totals, loaded
    // while the tile is ranked
    const uint32_t d_own = (uint32_t)tid;
    const uint32_t my_off = d_own < radix ? counts[(size_t)d_own * tiles + tile] : 0u;
    const uint32_t my_total = d_own < radix ? totals[d_own] : 0u;
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t idx = wbase + (uint32_t)k * 64u + (uint32_t)lane;
        const bool valid = idx < n;
        const uint32_t d = (key[k] >> shift) & dmask;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < bits; b++) {
            const bool set = (d >> b) & 1u;
            const uint64_t bal = __ballot(set);
            peers &= set ? bal : ~bal;
        }
        uint32_t r = 0;
        if (valid) {
            const uint32_t before = s_wcnt[wave][d];
            r = before + (uint32_t)__popcll(peers & lt);
            if ((uint32_t)(__ffsll((long long)peers) - 1) == (uint32_t)lane)
                s_wcnt[wave][d] = before + (uint32_t)__popcll(peers);
        }
        rnk[k] = r;
    }
    __syncthreads();

    // Per digit (lane d): offsets of each wave inside the tile and the tile
    // total; exclusive scans over digits of the tile totals (staging order)
    // and of the digit totals (global digit bases).
    uint32_t tile_count = 0;
    if (d_own < radix) {
        uint32_t run = 0;
        for (int w = 0; w < kSortWaves; w++) {
            const uint32_t c = s_wcnt[w][d_own];
            s_wcnt[w][d_own] = run;
            run += c;
        }
        tile_count = run;
    }
    uint32_t unused;
    const uint32_t tstart = block_exclusive_scan(tile_count, s_w, &unused);
    const uint32_t dbase = block_exclusive_scan(my_total, s_w, &unused);
    if (d_own < radix) {
        s_tstart[d_own] = tstart;
        s_gbase[d_own] = dbase + my_off;
    }
    __syncthreads();

    // Stage the tile in digit order.
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const uint32_t idx = wbase + (uint32_t)k * 64u + (uint32_t)lane;
        if (idx < n) {
            const uint32_t d = (key[k] >> shift) & dmask;
            const uint32_t pos = s_tstart[d] + s_wcnt[wave][d] + rnk[k];
            s_keys[pos] = key[k];
            s_vals[pos] = val[k];
        }
    }
    __syncthreads();
    const uint32_t count = (n - base) < (uint32_t)TILE ? (n - base) : (uint32_t)TILE;
    for (uint32_t q = tid; q < count; q += kSortThreads) {
        const uint32_t kk = s_keys[q];
        const uint32_t d = (kk >> shift) & dmask;
        const uint32_t dest = s_gbase[d] + (q - s_tstart[d]);
        if (dest < n) {
            kout[dest] = kk;
            vout[dest] = s_vals[q];
        }
    }
}

__global__ void k_copy_pairs(uint32_t* __restrict__ kd, uint32_t* __restrict__ vd, const uint32_t* __restrict__ ks,
                             const uint32_t* __restrict__ vs, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        kd[i] = ks[i];
        vd[i] = vs[i];
    }
}

__global__ void k_key_max(const uint32_t* __restrict__ keys, uint32_t n, uint32_t* __restrict__ out) {
    uint32_t m = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        m = max(m, keys[i]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// Persistent scratch for gSort/fg_sort_pairs.
struct SortScratch {
    void* ptr = nullptr;
    size_t bytes = 0;
};

static SortScratch& sort_scratch() {
    static SortScratch s;
    return s;
}

static void* ensure_scratch(size_t bytes) {
    SortScratch& s = sort_scratch();
    if (s.bytes >= bytes) return s.ptr;
    if (s.ptr) {
        (void)hipStreamSynchronize(state().stream);
        (void)hipFree(s.ptr);
        s.ptr = nullptr;
        s.bytes = 0;
    }
    size_t want = bytes + bytes / 8;
    if (!check_hip(hipMalloc(&s.ptr, want), "sort scratch hipMalloc")) return nullptr;
    s.bytes = want;
    return s.ptr;
}

static int run_sort(uint32_t* keys, uint32_t* vals, uint32_t n, int keyBits, void* ws, size_t ws_bytes,
                    bool iota, hipStream_t st) {
    if (n == 0) return FG_OK;
    if (n >= (1u << 30)) {
        set_error(FG_ERR_INVALID_ARGUMENT, "sort: n = %u exceeds 2^30 - 1", n);
        return FG_ERR_INVALID_ARGUMENT;
    }
    const SortPlan plan = make_plan(keyBits, state().sort_digit_bits);
    const SortLayout L = make_layout(n, keyBits);
    if (ws_bytes < L.total) {
        set_error(FG_ERR_INVALID_ARGUMENT, "sort: workspace %zu bytes < %zu needed", ws_bytes, L.total);
        return FG_ERR_INVALID_ARGUMENT;
    }
    char* w = static_cast<char*>(ws);
    uint32_t* totals = reinterpret_cast<uint32_t*>(w + L.totals);
    uint32_t* counts = reinterpret_cast<uint32_t*>(w + L.counts);
    uint32_t* t1k = reinterpret_cast<uint32_t*>(w + L.t1k);
    uint32_t* t1v = reinterpret_cast<uint32_t*>(w + L.t1v);
    uint32_t* t2k = plan.passes >= 3 ? reinterpret_cast<uint32_t*>(w + L.t2k) : nullptr;
    uint32_t* t2v = plan.passes >= 3 ? reinterpret_cast<uint32_t*>(w + L.t2v) : nullptr;
    const uint32_t tiles = sort_tiles(n);
    const bool small = sort_items(n) == kSortItemsSmall;

    const uint32_t* src_k = keys;
    const uint32_t* src_v = vals;
    for (int p = 0; p < plan.passes; p++) {
        uint32_t *dk, *dv;
        if (plan.passes == 1) {
            dk = t1k; dv = t1v;
        } else if (p == plan.passes - 1) {
            dk = keys; dv = vals;
        } else if (p % 2 == 0) {
            dk = t1k; dv = t1v;
        } else {
            dk = t2k; dv = t2v;
        }
        hipLaunchKernelGGL(small ? k_sort_upsweep<kSortItemsSmall> : k_sort_upsweep<kSortItemsLarge>, dim3(tiles),
                           dim3(kSortThreads), 0, st, src_k, n, plan.shift[p], plan.bits[p], counts, tiles);
        hipLaunchKernelGGL(k_sort_scan, dim3(1u << plan.bits[p]), dim3(kSortThreads), 0, st, counts, tiles, totals);
        hipLaunchKernelGGL(small ? k_sort_pass<kSortItemsSmall> : k_sort_pass<kSortItemsLarge>, dim3(tiles),
                           dim3(kSortThreads), 0, st, src_k, src_v, dk, dv, n, plan.shift[p], plan.bits[p], counts,
                           tiles, totals, (p == 0 && iota) ? 1 : 0);
        src_k = dk;
        src_v = dv;
    }
    if (plan.passes == 1)
        hipLaunchKernelGGL(k_copy_pairs, dim3(grid_for(n, 256)), dim3(256), 0, st, keys, vals, t1k, t1v, n);
    return check_launch("radix sort") ? FG_OK : fg_peek_error();
}

// ---------------------------------------------------------------------------
// Allocation-free sort: the workspace is a file-scope device array, one copy
// per GPU, kept until the process ends.  Host code never passes `scratch`
// itself to a kernel or copy (in host code the name is a host-side stand-in,
// not a device address): each device's copy is found once with
// hipGetSymbolAddress, with that device current, and only that address is
// used.
// ---------------------------------------------------------------------------
constexpr uint32_t kNoAllocMaxPairs = FG_SORT_NOALLOC_MAX_PAIRS;
constexpr size_t kNoAllocScratchBytes = FG_SORT_NOALLOC_SCRATCH_BYTES;
static_assert(kNoAllocScratchBytes % 256 == 0, "scratch size");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, or null (with the error recorded).
static void* noalloc_scratch() {
    int dev = -1;
    if (!check_hip(hipGetDevice(&dev), "fg_sort_pairs_noalloc hipGetDevice")) return nullptr;
    if (dev < 0 || dev >= kMaxDevices) {
        set_error(FG_ERR_UNSUPPORTED, "fg_sort_pairs_noalloc: device %d beyond %d", dev, kMaxDevices);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        if (!check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "fg_sort_pairs_noalloc hipGetSymbolAddress"))
            return nullptr;
        if (!ptr) {
            set_error(FG_ERR_HIP, "fg_sort_pairs_noalloc: no device address for the scratch array");
            return nullptr;
        }
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

}  // namespace fg

using namespace fg;

extern "C" {

int fg_sort_pairs_noalloc(unsigned int* keys, unsigned int* values, unsigned int n, int keyBits,
                          int values_are_iota) {
    if (n == 0) return FG_OK;
    if (!keys || !values) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sort_pairs_noalloc: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (keyBits < 1 || keyBits > 32) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sort_pairs_noalloc: keyBits %d outside 1..32", keyBits);
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (n > kNoAllocMaxPairs) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sort_pairs_noalloc: n = %u exceeds the %u pairs its %zu-byte "
                  "static workspace holds; use fg_sort_pairs_ws", n, kNoAllocMaxPairs, kNoAllocScratchBytes);
        return FG_ERR_INVALID_ARGUMENT;
    }
    const size_t need = make_layout(n, keyBits).total;
    if (need > kNoAllocScratchBytes) {  // cannot happen for n <= kNoAllocMaxPairs; kept as a guard
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sort_pairs_noalloc: %zu workspace bytes needed, %zu held", need,
                  kNoAllocScratchBytes);
        return FG_ERR_INVALID_ARGUMENT;
    }
    void* ws = noalloc_scratch();
    if (!ws) return fg_peek_error();
    return run_sort(keys, values, n, keyBits, ws, kNoAllocScratchBytes, values_are_iota != 0, state().stream);
}

size_t fg_sort_workspace_size(unsigned int n, int keyBits) { return make_layout(n, keyBits).total; }

int fg_sort_pairs_ws(unsigned int* keys, unsigned int* values, unsigned int n, int keyBits, void* workspace,
                     size_t workspace_bytes, int values_are_iota) {
    if (n && (!keys || !values || !workspace)) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sort_pairs_ws: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (keyBits < 1 || keyBits > 32) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sort_pairs_ws: keyBits %d outside 1..32", keyBits);
        return FG_ERR_INVALID_ARGUMENT;
    }
    return run_sort(keys, values, n, keyBits, workspace, workspace_bytes, values_are_iota != 0, state().stream);
}

static int sort_with_scratch(unsigned int* keys, unsigned int* values, unsigned int n, int keyBits, bool iota) {
    if (n == 0) return FG_OK;
    if (!keys || !values) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sort_pairs: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (keyBits < 1 || keyBits > 32) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_sort_pairs: keyBits %d outside 1..32", keyBits);
        return FG_ERR_INVALID_ARGUMENT;
    }
    const size_t need = fg_sort_workspace_size(n, keyBits);
    void* ws = ensure_scratch(need);
    if (!ws) return fg_peek_error();
    return run_sort(keys, values, n, keyBits, ws, sort_scratch().bytes, iota, state().stream);
}

int fg_sort_pairs(unsigned int* keys, unsigned int* values, unsigned int n, int keyBits) {
    return sort_with_scratch(keys, values, n, keyBits, false);
}

int fg_sort_pairs_iota(unsigned int* keys, unsigned int* values, unsigned int n, int keyBits) {
    return sort_with_scratch(keys, values, n, keyBits, true);
}

// K_FluidSim.cu:685-691.  The key width is read back from the device (one
// 4-byte copy, which synchronises the stream, as thrust's call does).
void gSort(unsigned int* _key, unsigned int* _values, int _numParticles) {
    if (_numParticles <= 0) return;
    if (!_key || !_values) {
        set_error(FG_ERR_INVALID_ARGUMENT, "gSort: null pointer");
        return;
    }
    const uint32_t n = (uint32_t)_numParticles;
    hipStream_t st = state().stream;
    void* ws = ensure_scratch(fg_sort_workspace_size(n, 32));
    if (!ws) return;
    uint32_t* kmax = reinterpret_cast<uint32_t*>(static_cast<char*>(ws) + make_layout(n, 32).counters) +
                     kMaxPasses;
    launch_fill_u32(kmax, 0u, 1, st);
    unsigned blocks = grid_for(n, 256 * 16);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_key_max, dim3(blocks), dim3(256), 0, st, _key, n, kmax);
    uint32_t host_max = 0;
    if (!check_hip(hipMemcpyAsync(&host_max, kmax, 4, hipMemcpyDeviceToHost, st), "gSort readback") ||
        !check_hip(hipStreamSynchronize(st), "gSort sync"))
        return;
    int bits = 1;
    while (bits < 32 && (host_max >> bits) != 0) bits++;
    run_sort(_key, _values, n, bits, ws, sort_scratch().bytes, false, st);
}

}  // extern "C"
