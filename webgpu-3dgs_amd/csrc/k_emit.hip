// k_emit.hip — emit one (tile, depth) key + splat id per overlapped tile.
//
// Result identical to emit_tile_refs.wgsl:56-79 (reference repo): a valid
// splat writes keys[offsets[id] + r] for r = 0..w*h-1 in row-major tile order,
// key = (ty*tiles_x + tx) << 32 | bits(-depth), value = id. Instead of one
// thread looping over one splat's tiles (divergent, uncoalesced), a workgroup
// owns 256 consecutive splats and therefore one contiguous output range
// [offsets[i0], offsets[i0+256]); thread j of the range finds its splat by a
// direct lookup in an LDS owner table (one byte store per key to build), so
// every store of keys and ids is coalesced. In the frame the splats come in
// depth order with per-position records from k_scan_rects (first key, first
// tile, width, depth key), read coalesced; with offsets instead, the tile
// rectangle and depth are read from the projected splats.
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
// contiguous key range with every store coalesced. The range is walked in
// windows of kOwnerWindow keys: each splat first marks the window slots it owns
// (owner[x] = local splat index), then thread j reads its owner directly.
constexpr uint32_t kOwnerWindow = 4096;

__device__ __forceinline__ void emit_chunk(const EmitArgs& a, uint32_t i0, uint32_t* s_start,
                                           uint32_t* s_tile0, uint32_t* s_width, uint32_t* s_keylo,
                                           uint32_t* s_sid, uint8_t* s_owner, uint32_t* s_hist) {
    const int tid = threadIdx.x;
    const uint32_t i = i0 + tid;
    const uint32_t last = min(i0 + kEmitThreads, a.n);
    const uint32_t radix = 1u << a.digit_bits;
    // first key of position p: the records' start column or the offsets
    auto first_key = [&](uint32_t p) { return a.recs ? a.recs[p].x : a.offsets[p]; };
    const uint32_t block_start = first_key(i0);
    const uint32_t block_end = first_key(last);
    uint32_t my_start = 0, my_count = 0, sid = i;
    if (i < a.n) {
        if (a.order) sid = a.order[i];
        if (a.recs) {
            // (first key, first tile, width, depth key) written by the scan
            const uint4 rec = a.recs[i];
            my_count = a.recs[i + 1].x - rec.x;
            my_start = rec.x - block_start;
            s_tile0[tid] = rec.y;
            s_width[tid] = rec.z;
            s_keylo[tid] = rec.w;
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
                s_tile0[tid] = t0y * a.tiles_x + t0x;
                s_width[tid] = t1x - t0x + 1u;
                s_keylo[tid] = __float_as_uint(-f0.z);
            }
        }
    }
    s_start[tid] = my_start;
    s_sid[tid] = sid;

    const uint32_t total = block_end - block_start;
    for (uint32_t w0 = 0; w0 < total; w0 += kOwnerWindow) {
        const uint32_t wn = min(kOwnerWindow, total - w0);
        __syncthreads();  // previous window's owners consumed (and the staging above visible)
        const uint32_t lo = max(my_start, w0), hi = min(my_start + my_count, w0 + wn);
        for (uint32_t x = lo; x < hi; x++) s_owner[x - w0] = (uint8_t)tid;
        __syncthreads();
        const uint32_t rounds = (wn + kEmitThreads - 1) / kEmitThreads;
        for (uint32_t round = 0; round < rounds; round++) {
            const uint32_t jw = round * kEmitThreads + tid;
            const bool live = jw < wn;
            uint64_t key = 0;
            if (live) {
                const uint32_t o = s_owner[jw];
                const uint32_t r = w0 + jw - s_start[o];
                const uint32_t w = s_width[o];
                const uint32_t ry = r / w;
                const uint32_t rx = r - ry * w;
                const uint32_t tile = s_tile0[o] + ry * a.tiles_x + rx;
                key = ((uint64_t)tile << 32) | (uint64_t)s_keylo[o];
                const uint64_t pos = (uint64_t)block_start + w0 + jw;
                if (pos < a.capacity) {
                    if (a.tiles) a.tiles[pos] = tile;
                    else a.keys[pos] = key;
                    a.values[pos] = s_sid[o];
                }
            }
            if (a.hist_partial && live) {
                // plain LDS atomics: a wavefront's keys come from ~18 splats in
                // depth order, so their tile digits rarely collide
                for (uint32_t p = 0; p < a.n_passes; p++) {
                    const uint32_t d = (uint32_t)(key >> (a.hist_shift0 + p * a.digit_bits)) & (radix - 1u);
                    atomicAdd(&s_hist[p * radix + d], 1u);
                }
            }
        }
    }
    __syncthreads();  // LDS staging is reused by the next chunk
}

// Grid-stride over 256-splat chunks; the digit histograms accumulate in LDS
// across a workgroup's chunks and leave as one plain row per workgroup
// (hist_partial[block][pass][digit]); k_hist_reduce folds the rows.
__global__ __launch_bounds__(kEmitThreads) void k_emit(EmitArgs a) {
    __shared__ uint32_t s_start[kEmitThreads];
    __shared__ uint32_t s_tile0[kEmitThreads];
    __shared__ uint32_t s_width[kEmitThreads];
    __shared__ uint32_t s_keylo[kEmitThreads];
    __shared__ uint32_t s_sid[kEmitThreads];
    __shared__ uint8_t s_owner[kOwnerWindow];
    __shared__ uint32_t s_hist[kMaxPasses * 256];
    const uint32_t radix = 1u << a.digit_bits;
    const uint32_t hist_words = a.hist_partial ? a.n_passes * radix : 0u;
    for (uint32_t k = threadIdx.x; k < hist_words; k += kEmitThreads) s_hist[k] = 0u;
    const uint32_t chunks = (a.n + kEmitThreads - 1) / kEmitThreads;
    for (uint32_t c = blockIdx.x; c < chunks; c += gridDim.x)
        emit_chunk(a, c * kEmitThreads, s_start, s_tile0, s_width, s_keylo, s_sid, s_owner, s_hist);
    if (a.hist_partial) {
        __syncthreads();
        uint32_t* row = a.hist_partial + (size_t)blockIdx.x * hist_words;
        for (uint32_t k = threadIdx.x; k < hist_words; k += kEmitThreads) row[k] = s_hist[k];
    }
}

// hist[w] += sum over rows r of partial[r][w]. A workgroup folds a block of
// kReduceWords words x kReduceRows rows: four groups of 64 threads each sum
// 16 rows (independent coalesced loads), LDS adds the groups, and one atomic
// per word and block leaves.
constexpr uint32_t kReduceWords = 64, kReduceRows = 64;

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
// workgroup (grid-stride, kHistKeys keys per thread per step, LDS atomics).
constexpr uint32_t kHistKeys = 4;

template <typename KeyT>
__global__ __launch_bounds__(256) void k_key_histogram(const KeyT* __restrict__ keys,
                                                       const uint32_t* __restrict__ n_dev,
                                                       uint32_t n_max, uint32_t* __restrict__ partial,
                                                       uint32_t digit_bits, uint32_t n_passes,
                                                       uint32_t shift0) {
    __shared__ uint32_t s_hist[kMaxPasses * 256];
    const uint32_t radix = 1u << digit_bits;
    const uint32_t words = n_passes * radix;
    for (uint32_t k = threadIdx.x; k < words; k += blockDim.x) s_hist[k] = 0u;
    __syncthreads();
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
            for (uint32_t p = 0; p < n_passes; p++)
                atomicAdd(&s_hist[p * radix + ((uint32_t)(k[j] >> (shift0 + p * digit_bits)) & (radix - 1u))],
                          1u);
        }
    }
    __syncthreads();
    uint32_t* row = partial + (size_t)blockIdx.x * words;
    for (uint32_t k = threadIdx.x; k < words; k += blockDim.x) row[k] = s_hist[k];
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
    hipLaunchKernelGGL(k_emit, dim3(grid), dim3(kEmitThreads), 0, s, a);
    if (a.hist_partial && a.hist)
        launch_hist_reduce(a.hist_partial, grid, a.n_passes * (1u << a.digit_bits), a.hist, s);
    return hipGetLastError();
}

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
    if (key_bytes == 8) {
        hipLaunchKernelGGL(k_key_histogram<uint64_t>, dim3(grid), dim3(256), 0, s,
                           static_cast<const uint64_t*>(keys), n_dev, n_max, hist_partial, digit_bits,
                           n_passes, shift0);
    } else if (key_bytes == 4) {
        hipLaunchKernelGGL(k_key_histogram<uint32_t>, dim3(grid), dim3(256), 0, s,
                           static_cast<const uint32_t*>(keys), n_dev, n_max, hist_partial, digit_bits,
                           n_passes, shift0);
    } else {
        return hipErrorInvalidValue;
    }
    launch_hist_reduce(hist_partial, grid, n_passes * (1u << digit_bits), hist, s);
    return hipGetLastError();
}

}  // namespace gs3d
