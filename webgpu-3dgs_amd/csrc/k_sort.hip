// k_sort.hip — one stable LSD radix pass over (u64 key, u32 value) pairs.
//
// The reference sorts with 8 passes x {per-workgroup LDS-atomic histogram,
// recursive Blelloch scan of a digit-major table, lane-0 serial ranking +
// scatter} (radix_histogram.wgsl:35-61, scan_local.wgsl, scan_add.wgsl,
// radix_scatter.wgsl:49-88; GaussianSplatRenderer.ts:1216-1247). Each pass is a
// stable counting sort on one digit, so the final order is (key, input index).
//
// Here a pass is ONE kernel (a "onesweep" pass):
//   1. a ticket gives the workgroup its tile (64 x ITEMS keys per wavefront:
//      16,384 keys per 1,024-thread workgroup at 16 keys per thread, 8,192 at
//      8) in launch order; keys and values are loaded together;
//   2. each wavefront ranks its keys with a wave-level multisplit: per digit
//      bit a 64-lane ballot narrows the mask of lanes holding the same digit;
//      the count of those lanes below this one = stable rank, the lowest lane
//      bumps the wave's private LDS histogram;
//   3. per digit: the tile's count is published as an aggregate, the workgroup
//      looks back over predecessor tiles' status words (decoupled lookback,
//      up to 4 threads per digit reading 8 predecessors each, 4 in the 8-key
//      shapes) for its exclusive prefix, then publishes the inclusive prefix;
//      the digit's global start comes from the digit histograms fused into
//      emit;
//   4. keys, then values, are placed digit-sorted in LDS and written out in
//      LDS order, so runs of equal digits leave as contiguous stores.
// HBM traffic per pass: 12 B read + 12 B written per 64-bit key (8 + 8 for
// 32-bit keys, 6 + 6 for 16-bit keys). Tiles beyond
// n = *n_dev exit at once, so a graph-captured launch can size the grid to the
// buffer capacity while the key count lives on the device.
#include "gs3d_internal.h"

namespace gs3d {

namespace {

__device__ __forceinline__ uint64_t load_status(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Two exclusive scans over the block with one barrier; s_tmp holds
// 2 * (THREADS / 64) words and is used once per workgroup.
template <int THREADS>
__device__ __forceinline__ void block_exclusive_scan2(uint32_t x, uint32_t y, uint32_t* s_tmp, uint32_t& ex,
                                                      uint32_t& ey) {
    constexpr int WAVES = THREADS / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    uint32_t a = x, b = y;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t ta = __shfl_up(a, d, kWave);
        const uint32_t tb = __shfl_up(b, d, kWave);
        if (lane >= d) {
            a += ta;
            b += tb;
        }
    }
    if (lane == kWave - 1) {
        s_tmp[wave] = a;
        s_tmp[WAVES + wave] = b;
    }
    __syncthreads();
    // the totals of the wavefronts before this one: lane l < wave reads wave
    // l's, a butterfly over the first WAVES lanes sums them, lane 0 has it
    // (one LDS read per lane instead of all WAVES totals held in registers)
    uint32_t pa = 0, pb = 0;
    if (lane < wave) {
        pa = s_tmp[lane];
        pb = s_tmp[WAVES + lane];
    }
#pragma unroll
    for (int d = 1; d < WAVES; d <<= 1) {
        pa += __shfl_xor(pa, d, kWave);
        pb += __shfl_xor(pb, d, kWave);
    }
    ex = __builtin_amdgcn_readfirstlane(pa) + a - x;
    ey = __builtin_amdgcn_readfirstlane(pb) + b - y;
}

// Lookback: per round trip each of up to 4 threads per digit reads the status
// words of kLookbackEach predecessor tiles with independent loads (32 tiles
// per round trip at 256 digits, 1,024 threads and 8 words per thread), and
// spins only on a word that is not published yet. Per-tile timestamps of a
// 118.9M-key pass: most tiles find a prefix in the first round; the lookback
// costs about one status round trip plus two barriers per round (3.5 us of an
// 18.5 us 16,384-key tile). Status words are tile-major,
// status[tile * RADIX + d], so a tile publishes all its digits with one
// coalesced store.
constexpr int kLookbackEach = 8;

// THREADS x ITEMS keys per tile. LDS: the staging buffer (one KeyT per key;
// values reuse it as u32), the per-wavefront digit counts, whose memory the
// lookback's per-thread partial sums reuse, and one per-digit delta: global
// output index = delta[d] + position in the tile, both u32 and added modulo
// 2^32 (scatter_index). The delta alone may wrap below zero; the sum is the
// true index, which is below n <= 2^32 - 1, so key counts of 2^31 and more
// index correctly.
// AUX: a third 32-bit array moves with the keys. kAuxGather: aux_out[position]
// = aux_in[value] (a gather through the sorted values, overlapped with the key
// and value write-out). kAuxCarry: aux_in is read at the key's own index, with
// the values, and lands beside it, like a second value array (coalesced reads;
// a pass chain carries it through every pass).
constexpr int kAuxNone = 0, kAuxGather = 1, kAuxCarry = 2;

template <int BITS, typename KeyT, bool IOTA, int THREADS, int ITEMS, int AUX>
__global__ __launch_bounds__(THREADS) void k_radix_pass(
    const KeyT* __restrict__ kin, const uint32_t* __restrict__ vin, KeyT* __restrict__ kout,
    uint32_t* __restrict__ vout, const uint32_t* __restrict__ n_dev,
    const uint32_t* __restrict__ hist_pass, uint64_t* status, uint32_t* ticket, uint32_t n_max,
    uint32_t shift, uint32_t epoch_arg, const uint32_t* __restrict__ aux_in,
    uint32_t* __restrict__ aux_out, const uint32_t* __restrict__ epoch_base, uint32_t match_bits) {
    const uint32_t epoch = resolve_epoch(epoch_arg, epoch_base);
    constexpr int RADIX = 1 << BITS;
    constexpr int WAVES = THREADS / kWave;
    constexpr int TILE = THREADS * ITEMS;
    constexpr uint32_t MASK = RADIX - 1;
    static_assert(THREADS >= RADIX && TILE >= kSortTile, "one thread per digit; status sized for kSortTile");
    // lookback: LB_SUB threads per digit, each reading LB_EACH predecessors
    constexpr int LB_SUB = THREADS / RADIX < 4 ? THREADS / RADIX : 4;
    // the 8-key shapes read 4 words per thread and round: with the other
    // register savings they stay at or under 64 VGPRs, so two workgroups
    // share a CU (one's loads overlap the other's ranking)
    constexpr int LB_EACH = ITEMS >= 16 ? kLookbackEach : kLookbackEach / 2;
    constexpr int LB_SPAN = LB_SUB * LB_EACH;
    constexpr int kCountBytes = WAVES * RADIX * 4;
    static_assert(LB_SUB * RADIX * 8 <= kCountBytes, "lookback partials fit the count memory");
    // staging: one KeyT per key, and at least 4 B per key, since the values
    // reuse it as u32 (16-bit tile ids stage in the low half)
    __shared__ __align__(16) uint32_t s_stage_words[TILE * (sizeof(KeyT) > 4 ? sizeof(KeyT) : 4) / 4];
    KeyT* const s_stage = reinterpret_cast<KeyT*>(s_stage_words);
    __shared__ __align__(16) uint32_t s_shared[kCountBytes / 4];
    __shared__ uint32_t s_delta[RADIX];
    __shared__ uint32_t s_tmp[2 * WAVES];
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_lb_open[2];
    uint32_t (*s_wave_hist)[RADIX] = reinterpret_cast<uint32_t (*)[RADIX]>(s_shared);

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    if (tid == 0) {
        s_part = atomicAdd(ticket, 1u);
        s_lb_open[0] = 0u;
        s_lb_open[1] = 0u;
    }
    for (int k = tid; k < WAVES * RADIX; k += THREADS) s_shared[k] = 0u;
    __syncthreads();
    const uint32_t part = s_part;
    const uint32_t n = *n_dev;
    const size_t tile_base = (size_t)part * TILE;
    // uniform over the block: past the keys, or more keys than the buffers hold
    // (the host re-runs the frame with larger buffers in that case)
    if (tile_base >= n || n > n_max) return;
    // the digit's global start (same for every tile) is loaded with the keys
    const uint32_t hist_d = tid < RADIX ? hist_pass[tid] : 0u;
    // Positions in the tile are u32 (< TILE): item it of this thread is tile
    // position lpos + 64 it, valid below tile_n.
    const uint32_t tile_n = (uint32_t)min((size_t)TILE, (size_t)n - tile_base);
    auto in_tile = [&](uint32_t j) -> bool { return j < tile_n; };
    const uint32_t lpos = (uint32_t)(wave * (kWave * ITEMS) + lane);
    const KeyT* __restrict__ kin_t = kin + tile_base;
    // Keys, values (and a carried array) are loaded together, the values
    // arriving while the keys are ranked. The loads are branch-free (a
    // position past the tile's keys reads the tile's last element; its key is
    // then replaced by the maximum key), so the ranking waits for the keys
    // alone. Values of an IOTA pass are positions, formed where they are
    // staged.
    KeyT k[ITEMS];
    uint32_t rank[ITEMS], val[ITEMS], aux[ITEMS];
    const uint32_t* __restrict__ vin_t = vin + tile_base;
    const uint32_t* __restrict__ aux_t = aux_in + tile_base;
#pragma unroll
    for (int it = 0; it < ITEMS; it++) k[it] = kin_t[min(lpos + (uint32_t)(it * kWave), tile_n - 1u)];
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const uint32_t jc = min(lpos + (uint32_t)(it * kWave), tile_n - 1u);
        if constexpr (!IOTA) val[it] = vin_t[jc];
        if constexpr (AUX == kAuxCarry) aux[it] = aux_t[jc];
    }
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        if (!in_tile(lpos + (uint32_t)(it * kWave))) k[it] = (KeyT)~(KeyT)0;
    }

    // Wave-level multisplit: per item, BITS ballots narrow each lane's group of
    // equal digits; rank = prior count of the digit in this wavefront + lanes
    // of the group below this one; the group's lowest lane bumps the count.
    // Two items share each ballot step (match_digit2); their counts are then
    // taken in item order.
    static_assert(ITEMS % 2 == 0, "items are matched in pairs");
    const uint32_t mb = __builtin_amdgcn_readfirstlane(match_bits);
    auto count = [&](int it, uint64_t peers, bool valid, uint32_t d) {
        uint32_t r = 0;
        if (valid) {
            // peers in lanes below this one (v_mbcnt: no mask of lower lanes)
            const uint32_t below =
                __builtin_amdgcn_mbcnt_hi((uint32_t)(peers >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0u));
            const uint32_t prior = s_wave_hist[wave][d];
            r = prior + below;
            if (below == 0u) s_wave_hist[wave][d] = prior + (uint32_t)__popcll(peers);
        }
        rank[it] = r;
        __builtin_amdgcn_wave_barrier();
    };
#pragma unroll
    for (int it = 0; it < ITEMS; it += 2) {
        const uint32_t j0 = lpos + (uint32_t)(it * kWave);
        const bool v0 = in_tile(j0), v1 = in_tile(j0 + (uint32_t)kWave);
        const uint32_t d0 = (uint32_t)(k[it] >> shift) & MASK;
        const uint32_t d1 = (uint32_t)(k[it + 1] >> shift) & MASK;
        uint64_t p0, p1;
        if (mb >= (uint32_t)BITS)
            match_digit2<BITS, true>(d0, v0, d1, v1, mb, p0, p1);
        else
            match_digit2<BITS, false>(d0, v0, d1, v1, mb, p0, p1);
        count(it, p0, v0, d0);
        count(it + 1, p1, v1, d1);
    }
    __syncthreads();

    // Per digit: counts across waves -> exclusive wave offsets and the tile
    // count; publish the count; the tile-local digit start is added to the
    // wave offsets once the block scan has it.
    uint32_t tile_count = 0;
    if (tid < RADIX) {
#pragma unroll
        for (int w = 0; w < WAVES; w++) {
            const uint32_t c = s_wave_hist[w][tid];
            s_wave_hist[w][tid] = tile_count;
            tile_count += c;
        }
        store_status(&status[(size_t)part * RADIX + tid],
                     status_word(epoch, part == 0 ? kFlagPrefix : kFlagAggregate, tile_count));
    }
    uint32_t local_start, global_start;
    block_exclusive_scan2<THREADS>(tid < RADIX ? tile_count : 0u, hist_d, s_tmp, local_start, global_start);
    if (tid < RADIX) {
#pragma unroll
        for (int w = 0; w < WAVES; w++) s_wave_hist[w][tid] += local_start;
    }
    __syncthreads();

    // Position in the tile (tile-local, known before the lookback); then the
    // counts' memory is free for the lookback. Keys go to LDS now, so they
    // leave the registers while the lookback waits on predecessor tiles.
#pragma unroll
    for (int it = 0; it < ITEMS; it++) rank[it] += s_wave_hist[wave][(uint32_t)(k[it] >> shift) & MASK];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        if (in_tile(lpos + (uint32_t)(it * kWave))) s_stage[rank[it]] = k[it];
    }

    // Decoupled lookback, LB_SPAN predecessors per round trip: thread
    // (sub, d) with sub < LB_SUB folds predecessors q - sub*LB_EACH - i
    // (i < LB_EACH) of digit d, nearest first, up to and including the first
    // inclusive prefix; the digit's thread (sub 0) then folds the LB_SUB
    // partial sums in order and stops at the first that met a prefix. The
    // round repeats while a digit is open; tile 0 publishes a prefix, so every
    // digit closes by then. (One wavefront per sub and digit range: each load
    // instruction reads one tile's status words, 512 contiguous bytes.)
    uint32_t excl = 0;
    if (part != 0) {
        uint32_t* s_lb_sum = s_shared;                      // [LB_SUB][RADIX]
        uint32_t* s_lb_met = s_shared + LB_SUB * RADIX;     // [LB_SUB][RADIX]
        const uint32_t lb_d = (uint32_t)tid & MASK;
        const int lb_sub = tid / RADIX;
        bool closed = false;  // digit lb_d met a prefix (sub-0 threads)
        int64_t q = (int64_t)part - 1;  // nearest predecessor not yet folded in
        for (int round = 0;; round++) {
            if (lb_sub < LB_SUB) {
                const int64_t q0 = q - (int64_t)lb_sub * LB_EACH;
                uint64_t w[LB_EACH];
#pragma unroll
                for (int i = 0; i < LB_EACH; i++)
                    w[i] = (q0 - i >= 0) ? load_status(&status[(size_t)(q0 - i) * RADIX + lb_d]) : 0ull;
                uint32_t sum = 0;
                bool met = false;
#pragma unroll
                for (int i = 0; i < LB_EACH; i++) {
                    if (met || q0 - i < 0) continue;
                    uint64_t x = w[i];
                    uint32_t flag = ((uint32_t)(x >> 34) == epoch) ? (uint32_t)((x >> 32) & 3u) : 0u;
                    while (flag == 0) {
                        x = load_status(&status[(size_t)(q0 - i) * RADIX + lb_d]);
                        flag = ((uint32_t)(x >> 34) == epoch) ? (uint32_t)((x >> 32) & 3u) : 0u;
                    }
                    sum += (uint32_t)x;
                    met = flag == kFlagPrefix;
                }
                s_lb_sum[lb_sub * RADIX + lb_d] = sum;
                s_lb_met[lb_sub * RADIX + lb_d] = met ? 1u : 0u;
            }
            __syncthreads();
            // s_lb_open[round & 1] collects "some digit still open"; the
            // other slot is cleared for the next round (its readers are past
            // the barrier above)
            if (tid < RADIX) {
#pragma unroll
                for (int sub = 0; sub < LB_SUB; sub++) {
                    if (!closed) {
                        excl += s_lb_sum[sub * RADIX + tid];
                        closed = s_lb_met[sub * RADIX + tid] != 0u;
                    }
                }
                if (!closed) s_lb_open[round & 1] = 1u;
            }
            if (tid == 0) s_lb_open[(round + 1) & 1] = 0u;
            q -= LB_SPAN;
            __syncthreads();
            if (s_lb_open[round & 1] == 0u || q < 0) break;
        }
    }
    if (tid < RADIX) {
        if (part != 0)
            store_status(&status[(size_t)part * RADIX + tid],
                         status_word(epoch, kFlagPrefix, excl + tile_count));
        s_delta[tid] = scatter_delta(global_start, excl, local_start);
    }
    __syncthreads();

    if constexpr (AUX == kAuxGather) {
        // issued after the lookback: loads return in order, so an earlier
        // gather would hold up the lookback's status loads
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            if (in_tile(lpos + (uint32_t)(it * kWave))) aux[it] = aux_in[val[it]];
        }
    }
    // keys out in LDS order (runs of equal digits are contiguous stores); the
    // output index of each sorted position is kept for the values (and aux)
    uint32_t dst[ITEMS];  // below n
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const uint32_t j = (uint32_t)(it * THREADS + tid);
        if (in_tile(j)) {
            const KeyT key = s_stage[j];
            dst[it] = (uint32_t)scatter_index(s_delta[(uint32_t)(key >> shift) & MASK], j);
            kout[dst[it]] = key;
        }
    }
    __syncthreads();
    // values: same permutation, reusing the staging memory
    uint32_t* s_vals = reinterpret_cast<uint32_t*>(s_stage);
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const uint32_t j = lpos + (uint32_t)(it * kWave);
        if (in_tile(j)) s_vals[rank[it]] = IOTA ? (uint32_t)(tile_base + j) : val[it];
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        if (in_tile((uint32_t)(it * THREADS + tid))) vout[dst[it]] = s_vals[it * THREADS + tid];
    }
    if constexpr (AUX != kAuxNone) {
        __syncthreads();
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            if (in_tile(lpos + (uint32_t)(it * kWave))) s_vals[rank[it]] = aux[it];
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < ITEMS; it++) {
            if (in_tile((uint32_t)(it * THREADS + tid))) aux_out[dst[it]] = s_vals[it * THREADS + tid];
        }
    }
}

}  // namespace

uint32_t radix_status_tiles(uint32_t n_max) { return (n_max + kSortTile - 1) / kSortTile + 1; }

// Tile shapes (threads x keys per thread). 64-bit keys: 16,384-key tiles
// (145 KB of LDS, one workgroup per CU) by default, 8,192 (73 KB, two
// workgroups per CU) on request. 32-bit keys: 8,192 by default (the depth
// passes over the splats), 16,384 for the tile-id passes over the references
// (fewer tiles in the lookback chain; measured faster at every size), 4,096
// on request.
constexpr int kSort64Threads = 1024, kSort64Items = 16;
constexpr int kSort64SmallThreads = 512, kSort64SmallItems = 16;
constexpr int kSort32Threads = 1024, kSort32Items = 8;
constexpr int kSort32SmallThreads = 512, kSort32SmallItems = 8;
constexpr int kSort32LargeThreads = 1024, kSort32LargeItems = 16;

template <int BITS, typename KeyT, bool IOTA, int AUX, int THREADS, int ITEMS>
static void launch_pass(const void* kin, const uint32_t* vin, void* kout, uint32_t* vout,
                        const uint32_t* n_dev, uint32_t n_max, const uint32_t* hist_pass,
                        uint64_t* status, uint32_t* ticket, uint32_t shift, uint32_t epoch,
                        const uint32_t* aux_in, uint32_t* aux_out, hipStream_t s,
                        const uint32_t* epoch_base, uint32_t match_bits) {
    const uint32_t grid = (n_max + THREADS * ITEMS - 1) / (THREADS * ITEMS);
    hipLaunchKernelGGL((k_radix_pass<BITS, KeyT, IOTA, THREADS, ITEMS, AUX>), dim3(grid), dim3(THREADS),
                       0, s, static_cast<const KeyT*>(kin), vin, static_cast<KeyT*>(kout), vout, n_dev,
                       hist_pass, status, ticket, n_max, shift, epoch, aux_in, aux_out, epoch_base, match_bits);
}

template <int BITS, typename KeyT, bool IOTA, int AUX>
static void launch_shape(int shape, const void* kin, const uint32_t* vin, void* kout, uint32_t* vout,
                         const uint32_t* n_dev, uint32_t n_max, const uint32_t* hist_pass, uint64_t* status,
                         uint32_t* ticket, uint32_t shift, uint32_t epoch, const uint32_t* aux_in,
                         uint32_t* aux_out, hipStream_t s, const uint32_t* epoch_base, uint32_t match_bits) {
    // shape: -1 the half-size tile, 0 the default, 1 the double-size tile (32-bit keys only)
    if constexpr (sizeof(KeyT) == 8) {
        if (shape < 0)
            launch_pass<BITS, KeyT, IOTA, AUX, kSort64SmallThreads, kSort64SmallItems>(
                kin, vin, kout, vout, n_dev, n_max, hist_pass, status, ticket, shift, epoch, aux_in, aux_out, s,
                epoch_base, match_bits);
        else
            launch_pass<BITS, KeyT, IOTA, AUX, kSort64Threads, kSort64Items>(
                kin, vin, kout, vout, n_dev, n_max, hist_pass, status, ticket, shift, epoch, aux_in, aux_out, s,
                epoch_base, match_bits);
    } else {
        if (shape < 0) {
            launch_pass<BITS, KeyT, IOTA, AUX, kSort32SmallThreads, kSort32SmallItems>(
                kin, vin, kout, vout, n_dev, n_max, hist_pass, status, ticket, shift, epoch, aux_in, aux_out, s,
                epoch_base, match_bits);
        } else if (shape > 0) {
            // the 16,384-key shape has no aux variant (rejected by launch_radix_pass)
            if constexpr (AUX == kAuxNone)
                launch_pass<BITS, KeyT, IOTA, AUX, kSort32LargeThreads, kSort32LargeItems>(
                    kin, vin, kout, vout, n_dev, n_max, hist_pass, status, ticket, shift, epoch, aux_in, aux_out,
                    s, epoch_base, match_bits);
        } else {
            launch_pass<BITS, KeyT, IOTA, AUX, kSort32Threads, kSort32Items>(
                kin, vin, kout, vout, n_dev, n_max, hist_pass, status, ticket, shift, epoch, aux_in, aux_out, s,
                epoch_base, match_bits);
        }
    }
}

// Threads and keys per thread of a pass shape (see launch_radix_pass for the
// key_tile values).
static void pass_shape(uint32_t key_bytes, uint32_t key_tile, uint32_t n_max, uint32_t* threads, uint32_t* items) {
    if (key_bytes == 8) {
        const bool small = key_tile == 8192 || (key_tile == 0 && n_max >= kSortSmallTileFrom);
        *threads = small ? kSort64SmallThreads : kSort64Threads;
        *items = small ? kSort64SmallItems : kSort64Items;
    } else if (key_tile == 4096) {
        *threads = kSort32SmallThreads;
        *items = kSort32SmallItems;
    } else if (key_tile == 16384) {
        *threads = kSort32LargeThreads;
        *items = kSort32LargeItems;
    } else {
        *threads = kSort32Threads;
        *items = kSort32Items;
    }
}

LaunchShape radix_pass_launch_shape(uint32_t key_bytes, uint32_t key_tile, uint32_t n) {
    uint32_t threads = 0, items = 0;
    pass_shape(key_bytes, key_tile, n, &threads, &items);
    return {threads, ((uint64_t)n + threads * items - 1) / (threads * items)};
}

hipError_t launch_radix_pass(const void* kin, const uint32_t* vin, void* kout, uint32_t* vout,
                             uint32_t key_bytes, bool iota_values, const uint32_t* n_dev,
                             uint32_t n_max, const uint32_t* hist_pass, uint64_t* status,
                             uint32_t* ticket, uint32_t shift, uint32_t digit_bits, uint32_t epoch,
                             const uint32_t* aux_in, uint32_t* aux_out, hipStream_t s,
                             const uint32_t* epoch_base, uint32_t key_tile, uint32_t match_bits,
                             bool aux_carry) {
    if (n_max == 0) return hipSuccess;
    if (match_bits == 0 || match_bits > digit_bits) match_bits = digit_bits;
    // key_tile: keys per workgroup, 0 = the default shape for the key width
    // (16,384 for 64-bit keys, 8,192 for 32-bit keys); 64-bit: 8,192 or
    // 16,384; 32- and 16-bit: 4,096, 8,192 or 16,384 (16,384 not with aux)
    int shape;  // -1 the half-size tile, 0 the default, 1 the double-size tile
    if (key_bytes == 8) {
        if (key_tile != 0 && key_tile != 8192 && key_tile != 16384) return hipErrorInvalidValue;
        shape = (key_tile == 8192 || (key_tile == 0 && n_max >= kSortSmallTileFrom)) ? -1 : 0;
    } else {
        if (key_tile != 0 && key_tile != 4096 && key_tile != 8192 && key_tile != 16384) return hipErrorInvalidValue;
        shape = key_tile == 4096 ? -1 : (key_tile == 16384 ? 1 : 0);
    }
    const bool aux = aux_in != nullptr && aux_out != nullptr;
    // 16-bit keys (tile ids of frames up to 65,536 tiles) take the 32-bit
    // shapes; aux rides on 32-bit key passes of the 4,096- and 8,192-key shapes; the
    // gather also needs the values read from vin
    if (aux && (key_bytes != 4 || shape > 0 || (!aux_carry && iota_values))) return hipErrorInvalidValue;
    const int key = (digit_bits == 8 ? 100 : digit_bits == 4 ? 200 : 0) +
                    (key_bytes == 8 ? 10 : key_bytes == 4 ? 20 : key_bytes == 2 ? 30 : 0) +
                    (iota_values ? 1 : 0) + (aux ? (aux_carry ? 4 : 2) : 0);
#define GS3D_PASS(B, K, I, A)                                                                      \
    launch_shape<B, K, I, A>(shape, kin, vin, kout, vout, n_dev, n_max, hist_pass, status, ticket, shift, \
                             epoch, aux_in, aux_out, s, epoch_base, match_bits)
    switch (key) {
        case 110: GS3D_PASS(8, uint64_t, false, kAuxNone); break;
        case 111: GS3D_PASS(8, uint64_t, true, kAuxNone); break;
        case 120: GS3D_PASS(8, uint32_t, false, kAuxNone); break;
        case 121: GS3D_PASS(8, uint32_t, true, kAuxNone); break;
        case 122: GS3D_PASS(8, uint32_t, false, kAuxGather); break;
        case 124: GS3D_PASS(8, uint32_t, false, kAuxCarry); break;
        case 125: GS3D_PASS(8, uint32_t, true, kAuxCarry); break;
        case 130: GS3D_PASS(8, uint16_t, false, kAuxNone); break;
        case 210: GS3D_PASS(4, uint64_t, false, kAuxNone); break;
        case 211: GS3D_PASS(4, uint64_t, true, kAuxNone); break;
        case 220: GS3D_PASS(4, uint32_t, false, kAuxNone); break;
        case 221: GS3D_PASS(4, uint32_t, true, kAuxNone); break;
        case 222: GS3D_PASS(4, uint32_t, false, kAuxGather); break;
        case 224: GS3D_PASS(4, uint32_t, false, kAuxCarry); break;
        case 225: GS3D_PASS(4, uint32_t, true, kAuxCarry); break;
        case 230: GS3D_PASS(4, uint16_t, false, kAuxNone); break;
        default: return hipErrorInvalidValue;
    }
#undef GS3D_PASS
    return hipGetLastError();
}

namespace {
__global__ void k_bump_epoch(uint32_t* epoch_base) {
    if (threadIdx.x == 0) epoch_base[0] = (epoch_base[0] + kEpochStride) & kEpochMask;
}
}  // namespace

hipError_t launch_bump_epoch(uint32_t* epoch_base, hipStream_t s) {
    hipLaunchKernelGGL(k_bump_epoch, dim3(1), dim3(64), 0, s, epoch_base);
    return hipGetLastError();
}

}  // namespace gs3d
