// k_emit.hip — emit one (tile, depth) key + splat id per overlapped tile.
//
// Result identical to emit_tile_refs.wgsl:56-79 (reference repo): a valid
// splat writes keys[offsets[id] + r] for r = 0..w*h-1 in row-major tile order,
// key = (ty*tiles_x + tx) << 32 | bits(-depth), value = id. Instead of one
// thread looping over one splat's tiles (divergent, uncoalesced), a workgroup
// owns 256 consecutive splats and therefore one contiguous output range
// [offsets[i0], offsets[i0+256]); thread j of the range finds its splat from
// the splats' first slots marked in LDS (see emit_chunk), so every store of
// keys and ids is coalesced and no thread loops over a splat's tiles. In the
// frame the splats come in depth order with per-position records from
// k_scan_rects (first key, first tile, width, depth key), read coalesced; with
// offsets instead, the tile rectangle and depth are read from the projected
// splats.
//
// With `tiles` set, only the tile id is stored (4 B instead of 8): the stream
// is emitted in stable depth order, so sorting the tile ids alone finishes
// the (tile, depth) order; the 64-bit keys are rebuilt only for readback.
//
// Fused here: the digit histograms of every radix pass that the sort will run,
// so the sort does not re-read the keys to count them. Keys past `capacity`
// are counted but not stored; the host sees that in ref_counter and re-runs.
#include "gs3d_internal.h"

namespace gs3d {

namespace {

constexpr int kEmitThreads = 256;

// One 256-splat chunk: stage the chunk's splats in LDS, then emit its
// contiguous key range. The range is walked in windows of kOwnerWindow keys.
// Each splat that overlaps a window marks only its first slot there (owner
// byte + one bit of a slot mask; the ranges are disjoint and slot 0 always
// belongs to some splat); one wavefront then max-scans the mask words, so a
// slot's splat is the last mark at or before it, found in O(1) whatever the
// splats' sizes. A thread takes kSlotsPerThread consecutive slots: one lookup
// and one division for the first, then the next slot is the next tile of the
// same row, the first tile of its next row, or the first tile of the splat
// that starts there (its mask bit). Two slots per thread measured fastest
// (halves the dependent LDS reads per key; with four, the stores of a
// wavefront stride 16 B and the kernel slows down).
constexpr uint32_t kOwnerWindow = 4096;
constexpr uint32_t kSlotsPerThread = 2;
constexpr uint32_t kMarkWords = kOwnerWindow / 32;
static_assert(kEmitThreads >= kWave + kMarkWords, "mask clearing needs its own threads");
static_assert(kMarkWords == 2 * kWave, "the mask scan gives each lane of one wavefront two words");
static_assert(32 % kSlotsPerThread == 0, "a thread's slots share one mask word");

struct EmitShared {
    uint32_t start[kEmitThreads];  // first slot of the splat, relative to the chunk's first key
    uint32_t tile0[kEmitThreads];  // tile id of the splat's top-left tile
    uint32_t width[kEmitThreads];  // rectangle width in tiles
    uint32_t keylo[kEmitThreads];  // depth key
    uint32_t sid[kEmitThreads];    // splat id
    uint32_t mark[2][kMarkWords];  // slot masks, alternating; the one a window uses starts clear
    int32_t prev[kMarkWords];      // last marked slot before each mask word (-1: none)
    uint8_t owner[kOwnerWindow];   // at marked slots: the chunk-local splat
};

// The digit histograms (emit and k_key_histogram) are kept in LDS in
// kHistCopies copies, copy = thread % kHistCopies, interleaved word by word:
// lanes of a wavefront that share a digit (most of them, for high digits) add
// to kHistCopies different banks instead of one word. The copies are summed
// when the row is written.
constexpr uint32_t kHistCopies = 4;

__device__ __forceinline__ void emit_chunk(const EmitArgs& a, uint32_t i0, EmitShared& s, uint32_t* s_hist,
                                           uint32_t& cur) {
    const int tid = threadIdx.x;
    const uint32_t i = i0 + tid;
    const uint32_t last = min(i0 + kEmitThreads, a.n);
    // first key of position p: the records' start column or the offsets
    auto first_key = [&](uint32_t p) { return a.recs ? a.recs[p].x : a.offsets[p]; };
    const uint32_t block_start = first_key(i0);
    const uint32_t block_end = first_key(last);
    uint32_t my_start = 0, my_count = 0, sid = i, width = 1;
    if (i < a.n) {
        if (a.order) sid = a.order[i];
        if (a.recs) {
            // (first key, first tile, width, depth key) written by the scan
            const uint4 rec = a.recs[i];
            my_count = a.recs[i + 1].x - rec.x;
            my_start = rec.x - block_start;
            s.tile0[tid] = rec.y;
            width = rec.z;
            s.keylo[tid] = rec.w;
        } else {
            const uint32_t off = a.offsets[i];
            my_count = a.offsets[i + 1] - off;  // > 0 exactly for valid splats
            my_start = off - block_start;
            if (my_count) {
                const float4* ps = reinterpret_cast<const float4*>(a.projected + sid);
                const float4 f0 = ps[0];
                const float4 f3 = ps[3];
                const uint32_t t0x = __float_as_uint(f3.x), t0y = __float_as_uint(f3.y);
                const uint32_t t1x = __float_as_uint(f3.z);
                s.tile0[tid] = t0y * a.tiles_x + t0x;
                width = t1x - t0x + 1u;
                s.keylo[tid] = __float_as_uint(-f0.z);
            }
        }
    }
    if (my_count == 0) width = 1;
    s.width[tid] = width;
    s.start[tid] = my_start;
    s.sid[tid] = sid;

    const uint32_t total = block_end - block_start;
    for (uint32_t w0 = 0; w0 < total; w0 += kOwnerWindow) {
        const uint32_t wn = min(kOwnerWindow, total - w0);
        __syncthreads();  // previous window's slots consumed (and the staging above visible)
        const uint32_t lo = max(my_start, w0), hi = min(my_start + my_count, w0 + wn);
        if (lo < hi) {
            const uint32_t x = lo - w0;
            s.owner[x] = (uint8_t)tid;
            atomicOr(&s.mark[cur][x >> 5], 1u << (x & 31u));
        }
        __syncthreads();
        if (tid < kWave) {
            // lane l: mask words 2l and 2l+1; inclusive max-scan of their last marks
            const uint32_t ma = s.mark[cur][2 * tid], mb = s.mark[cur][2 * tid + 1];
            const int32_t la = ma ? (int32_t)(64 * tid + 31 - __clz((int)ma)) : -1;
            const int32_t lb = mb ? (int32_t)(64 * tid + 63 - __clz((int)mb)) : -1;
            int32_t v = la > lb ? la : lb;
#pragma unroll
            for (int d = 1; d < kWave; d <<= 1) {
                const int32_t t = __shfl_up(v, d, kWave);
                if (tid >= d) v = t > v ? t : v;
            }
            int32_t ex = __shfl_up(v, 1, kWave);
            if (tid == 0) ex = -1;
            s.prev[2 * tid] = ex;
            s.prev[2 * tid + 1] = la > ex ? la : ex;
        } else if (tid < kWave + (int)kMarkWords) {
            s.mark[cur ^ 1u][tid - kWave] = 0u;  // the next window's mask
        }
        __syncthreads();
        constexpr uint32_t SPT = kSlotsPerThread;
        const uint32_t rounds = (wn + kEmitThreads * SPT - 1) / (kEmitThreads * SPT);
        const uint32_t radix = 1u << a.digit_bits;
        for (uint32_t round = 0; round < rounds; round++) {
            const uint32_t g0 = (round * kEmitThreads + tid) * SPT;
            if (g0 < wn) {
                const uint32_t word = g0 >> 5;
                const uint32_t mword = s.mark[cur][word];
                const uint32_t m = mword & ((2u << (g0 & 31u)) - 1u);
                const int32_t at = m ? (int32_t)((word << 5) + 31 - __clz((int)m)) : s.prev[word];
                uint32_t o = s.owner[at];
                const uint32_t r = w0 + g0 - s.start[o];
                uint32_t w = s.width[o];
                const uint32_t ry = r / w;
                uint32_t rx = r - ry * w;
                uint32_t tile = s.tile0[o] + ry * a.tiles_x + rx;
                uint32_t keylo = s.keylo[o], sid = s.sid[o];
#pragma unroll
                for (uint32_t k = 0; k < SPT; k++) {
                    const uint32_t slot = g0 + k;
                    if (slot >= wn) break;
                    if (k > 0) {
                        if ((mword >> (slot & 31u)) & 1u) {
                            // the next splat starts at this slot
                            o = s.owner[slot];
                            w = s.width[o];
                            tile = s.tile0[o];
                            keylo = s.keylo[o];
                            sid = s.sid[o];
                            rx = 0;
                        } else if (++rx == w) {
                            rx = 0;
                            tile += a.tiles_x - w + 1u;
                        } else {
                            tile++;
                        }
                    }
                    const uint64_t key = ((uint64_t)tile << 32) | (uint64_t)keylo;
                    const uint64_t pos = (uint64_t)block_start + w0 + slot;
                    if (pos < a.capacity) {
                        if (a.tiles16) a.tiles16[pos] = (uint16_t)tile;
                        else if (a.tiles) a.tiles[pos] = tile;
                        else a.keys[pos] = key;
                        a.values[pos] = sid;
                    }
                    if (a.hist_partial) {
                        for (uint32_t p = 0; p < a.n_passes; p++) {
                            const uint32_t d = (uint32_t)(key >> (a.hist_shift0 + p * a.digit_bits)) & (radix - 1u);
                            atomicAdd(&s_hist[(p * radix + d) * kHistCopies + tid % kHistCopies], 1u);
                        }
                    }
                }
            }
        }
        cur ^= 1u;
    }
    __syncthreads();  // LDS staging is reused by the next chunk
}

// Grid-stride over 256-splat chunks; the digit histograms accumulate in LDS
// across a workgroup's chunks and leave as one plain row per workgroup
// (hist_partial[block][pass][digit]); k_hist_reduce folds the rows.
__global__ __launch_bounds__(kEmitThreads) void k_emit(EmitArgs a) {
    __shared__ EmitShared s;
    extern __shared__ uint32_t s_hist[];  // [pass * radix + digit][copy], launch-sized
    const uint32_t radix = 1u << a.digit_bits;
    const uint32_t hist_words = a.hist_partial ? a.n_passes * radix : 0u;
    for (uint32_t k = threadIdx.x; k < hist_words * kHistCopies; k += kEmitThreads) s_hist[k] = 0u;
    for (uint32_t k = threadIdx.x; k < 2 * kMarkWords; k += kEmitThreads) (&s.mark[0][0])[k] = 0u;
    uint32_t cur = 0;
    const uint32_t chunks = (a.n + kEmitThreads - 1) / kEmitThreads;
    for (uint32_t c = blockIdx.x; c < chunks; c += gridDim.x) emit_chunk(a, c * kEmitThreads, s, s_hist, cur);
    if (a.hist_partial) {
        __syncthreads();
        uint32_t* row = a.hist_partial + (size_t)blockIdx.x * hist_words;
        for (uint32_t k = threadIdx.x; k < hist_words; k += kEmitThreads) {
            uint32_t sum = 0;
#pragma unroll
            for (uint32_t c = 0; c < kHistCopies; c++) sum += s_hist[k * kHistCopies + c];
            row[k] = sum;
        }
    }
}

// hist[w] += sum over rows r of partial[r][w]. A workgroup folds a block of
// kReduceWords words x kReduceRows rows: four groups of 64 threads each sum
// 8 rows (independent coalesced loads), LDS adds the groups, and one atomic
// per word and block leaves. Rows per workgroup, measured on the 1M frame:
// 8 / 16 / 32 / 64 / 128 / 256 -> 0.425 / 0.424 / 0.424 / 0.427 / 0.438 /
// 0.457 ms.
constexpr uint32_t kReduceWords = 64, kReduceRows = 32;
static_assert(kReduceWords * 4 == 256 && kReduceRows % 4 == 0,
              "k_hist_reduce: four groups of kReduceWords threads, a quarter of the rows each");

__global__ __launch_bounds__(256) void k_hist_reduce(const uint32_t* __restrict__ partial,
                                                     uint32_t rows, uint32_t words,
                                                     uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_sum[3][kReduceWords];
    const uint32_t wl = threadIdx.x % kReduceWords, g = threadIdx.x / kReduceWords;
    const uint32_t w = blockIdx.x * kReduceWords + wl;
    const uint32_t r0 = blockIdx.y * kReduceRows + g * (kReduceRows / 4);
    uint32_t sum = 0;
    if (w < words) {
#pragma unroll
        for (uint32_t i = 0; i < kReduceRows / 4; i++) {
            const uint32_t r = r0 + i;
            if (r < rows) sum += partial[(size_t)r * words + w];
        }
    }
    if (g > 0) s_sum[g - 1][wl] = sum;
    __syncthreads();
    if (g == 0 && w < words) {
        sum += s_sum[0][wl] + s_sum[1][wl] + s_sum[2][wl];
        if (sum) atomicAdd(&hist[w], sum);
    }
}

void launch_hist_reduce(const uint32_t* partial, uint32_t rows, uint32_t words, uint32_t* hist,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_hist_reduce, dim3((words + kReduceWords - 1) / kReduceWords,
                                           (rows + kReduceRows - 1) / kReduceRows),
                       dim3(256), 0, s, partial, rows, words, hist);
}

// Digit histograms of n = *n_dev keys for every pass, one plain row per
// workgroup (grid-stride, kHistKeys keys per thread per step, LDS atomics into
// kHistCopies copies).
constexpr uint32_t kHistKeys = 4;

template <typename KeyT>
__global__ __launch_bounds__(256) void k_key_histogram(const KeyT* __restrict__ keys,
                                                       const uint32_t* __restrict__ n_dev,
                                                       uint32_t n_max, uint32_t* __restrict__ partial,
                                                       uint32_t digit_bits, uint32_t n_passes,
                                                       uint32_t shift0) {
    extern __shared__ uint32_t s_hist[];  // [pass * radix + digit][copy]
    const uint32_t radix = 1u << digit_bits;
    const uint32_t words = n_passes * radix;
    for (uint32_t k = threadIdx.x; k < words * kHistCopies; k += blockDim.x) s_hist[k] = 0u;
    __syncthreads();
    const uint32_t copy = threadIdx.x % kHistCopies;
    uint32_t n = *n_dev;
    if (n > n_max) n = 0;  // keys did not fit: count nothing, the frame is re-run
    const size_t stride = (size_t)gridDim.x * blockDim.x * kHistKeys;
    for (size_t base = (size_t)blockIdx.x * blockDim.x * kHistKeys; base < n; base += stride) {
        KeyT k[kHistKeys];
#pragma unroll
        for (uint32_t j = 0; j < kHistKeys; j++) {
            const size_t i = base + j * blockDim.x + threadIdx.x;
            k[j] = i < n ? keys[i] : (KeyT)0;
        }
#pragma unroll
        for (uint32_t j = 0; j < kHistKeys; j++) {
            if (base + j * blockDim.x + threadIdx.x >= n) continue;
            for (uint32_t p = 0; p < n_passes; p++) {
                const uint32_t w = p * radix + ((uint32_t)(k[j] >> (shift0 + p * digit_bits)) & (radix - 1u));
                atomicAdd(&s_hist[w * kHistCopies + copy], 1u);
            }
        }
    }
    __syncthreads();
    uint32_t* row = partial + (size_t)blockIdx.x * words;
    for (uint32_t k = threadIdx.x; k < words; k += blockDim.x) {
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t c = 0; c < kHistCopies; c++) sum += s_hist[k * kHistCopies + c];
        row[k] = sum;
    }
}

}  // namespace

uint32_t emit_grid(uint32_t n) {
    const uint32_t chunks = (n + kEmitThreads - 1) / kEmitThreads;
    return chunks < kEmitMaxBlocks ? (chunks ? chunks : 1u) : kEmitMaxBlocks;
}

hipError_t launch_emit(const EmitArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    if (a.n_passes > kMaxPasses || (a.digit_bits != 4 && a.digit_bits != 8)) return hipErrorInvalidValue;
    const uint32_t grid = emit_grid(a.n);
    const size_t lds = a.hist_partial ? (size_t)a.n_passes * (1u << a.digit_bits) * kHistCopies * 4u : 0u;
    hipLaunchKernelGGL(k_emit, dim3(grid), dim3(kEmitThreads), lds, s, a);
    if (a.hist_partial && a.hist)
        launch_hist_reduce(a.hist_partial, grid, a.n_passes * (1u << a.digit_bits), a.hist, s);
    return hipGetLastError();
}

LaunchShape emit_launch_shape(uint32_t n) { return {(uint32_t)kEmitThreads, emit_grid(n)}; }
LaunchShape key_histogram_launch_shape(uint32_t n) { return {256u, key_hist_rows(n)}; }

uint32_t key_hist_rows(uint32_t n_max) {
    const uint32_t blocks = (n_max + 256u * kHistKeys - 1u) / (256u * kHistKeys);
    return blocks < 1024u ? (blocks ? blocks : 1u) : 1024u;
}

hipError_t launch_key_histogram(const void* keys, uint32_t key_bytes, const uint32_t* n_dev,
                                uint32_t n_max, uint32_t* hist, uint32_t* hist_partial,
                                uint32_t digit_bits, uint32_t n_passes, uint32_t shift0,
                                hipStream_t s) {
    if (n_max == 0 || n_passes == 0) return hipSuccess;
    if (n_passes > kMaxPasses || (digit_bits != 4 && digit_bits != 8)) return hipErrorInvalidValue;
    const uint32_t grid = key_hist_rows(n_max);
    const size_t lds = (size_t)n_passes * (1u << digit_bits) * kHistCopies * 4u;  // <= 64 KB
    if (key_bytes == 8) {
        hipLaunchKernelGGL(k_key_histogram<uint64_t>, dim3(grid), dim3(256), lds, s,
                           static_cast<const uint64_t*>(keys), n_dev, n_max, hist_partial, digit_bits,
                           n_passes, shift0);
    } else if (key_bytes == 4) {
        hipLaunchKernelGGL(k_key_histogram<uint32_t>, dim3(grid), dim3(256), lds, s,
                           static_cast<const uint32_t*>(keys), n_dev, n_max, hist_partial, digit_bits,
                           n_passes, shift0);
    } else {
        return hipErrorInvalidValue;
    }
    launch_hist_reduce(hist_partial, grid, n_passes * (1u << digit_bits), hist, s);
    return hipGetLastError();
}

}  // namespace gs3d
