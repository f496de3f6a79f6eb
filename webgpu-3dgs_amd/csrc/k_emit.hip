// k_emit.hip — emit one (tile, depth) key + splat id per overlapped tile.
//
// Result identical to emit_tile_refs.wgsl:56-79 (reference repo): a valid
// splat writes keys[offsets[id] + r] for r = 0..w*h-1 in row-major tile order,
// key = (ty*tiles_x + tx) << 32 | bits(-depth), value = id. Instead of one
// thread looping over one splat's tiles (divergent, uncoalesced), a workgroup
// owns 256 consecutive splats and therefore one contiguous output range
// [offsets[i0], offsets[i0+256]); thread j of the range finds its splat by a
// binary search over the 256 local starts in LDS, so every store is coalesced
// and a 6000-tile splat costs the same as 6000 small ones.
//
// Fused here: the digit histograms of every radix pass that the sort will run
// (LDS atomics per workgroup, one global atomic per non-zero bin), so the sort
// does not re-read the keys to count them. Keys past `capacity` are counted
// but not stored; the host detects that from ref_counter and re-emits.
#include "gs3d_internal.h"

namespace gs3d {

namespace {

constexpr int kEmitThreads = 256;

__global__ __launch_bounds__(kEmitThreads) void k_emit(EmitArgs a) {
    __shared__ uint32_t s_start[kEmitThreads];
    __shared__ uint32_t s_tmin_x[kEmitThreads];
    __shared__ uint32_t s_tmin_y[kEmitThreads];
    __shared__ uint32_t s_width[kEmitThreads];
    __shared__ uint32_t s_keylo[kEmitThreads];
    __shared__ uint32_t s_hist[kMaxPasses * 256];

    const int tid = threadIdx.x;
    const uint32_t i0 = blockIdx.x * kEmitThreads;
    const uint32_t i = i0 + tid;
    const uint32_t last = min(i0 + kEmitThreads, a.n);
    const uint32_t radix = 1u << a.digit_bits;
    const uint32_t hist_words = a.hist ? a.n_passes * radix : 0u;
    for (uint32_t k = tid; k < hist_words; k += kEmitThreads) s_hist[k] = 0u;

    const uint32_t block_start = a.offsets[i0];
    const uint32_t block_end = a.offsets[last];
    if (i < a.n) {
        const uint32_t off = a.offsets[i];
        const uint32_t cnt = a.offsets[i + 1] - off;  // > 0 exactly for valid splats
        s_start[tid] = off - block_start;
        if (cnt) {
            const float4* ps = reinterpret_cast<const float4*>(a.projected + i);
            const float4 f0 = ps[0];
            const float4 f3 = ps[3];
            const uint32_t t0x = __float_as_uint(f3.x), t0y = __float_as_uint(f3.y);
            const uint32_t t1x = __float_as_uint(f3.z);
            s_tmin_x[tid] = t0x;
            s_tmin_y[tid] = t0y;
            s_width[tid] = t1x - t0x + 1u;
            s_keylo[tid] = __float_as_uint(-f0.z);
        } else {
            s_width[tid] = 1u;
        }
    } else {
        s_start[tid] = block_end - block_start;  // past-the-end sentinel for the search
        s_width[tid] = 1u;
    }
    __syncthreads();

    const uint32_t total = block_end - block_start;
    const uint32_t nloc = last - i0;
    for (uint32_t j = tid; j < total; j += kEmitThreads) {
        // largest s with s_start[s] <= j (upper_bound - 1) over [0, nloc)
        uint32_t lo = 0, hi = nloc;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_start[mid] <= j) lo = mid; else hi = mid;
        }
        const uint32_t r = j - s_start[lo];
        const uint32_t w = s_width[lo];
        const uint32_t ry = r / w;
        const uint32_t rx = r - ry * w;
        const uint32_t tile = (s_tmin_y[lo] + ry) * a.tiles_x + s_tmin_x[lo] + rx;
        const uint64_t key = ((uint64_t)tile << 32) | (uint64_t)s_keylo[lo];
        const uint64_t pos = (uint64_t)block_start + j;
        if (pos < a.capacity) {
            a.keys[pos] = key;
            a.values[pos] = i0 + lo;
        }
        if (a.hist) {
            for (uint32_t p = 0; p < a.n_passes; p++) {
                const uint32_t d = (uint32_t)(key >> a.shifts[p]) & (radix - 1u);
                atomicAdd(&s_hist[p * radix + d], 1u);
            }
        }
    }
    if (a.hist) {
        __syncthreads();
        for (uint32_t k = tid; k < hist_words; k += kEmitThreads) {
            const uint32_t c = s_hist[k];
            if (c) atomicAdd(&a.hist[k], c);
        }
    }
}

// Digit histograms of n = *n_dev keys for every pass (standalone sort path).
__global__ __launch_bounds__(256) void k_digit_histogram(const uint64_t* keys, const uint32_t* n_dev,
                                                         uint32_t* hist, uint32_t digit_bits,
                                                         uint32_t n_passes, const uint32_t* shifts_dev) {
    __shared__ uint32_t s_hist[kMaxPasses * 256];
    __shared__ uint32_t s_shift[kMaxPasses];
    const uint32_t radix = 1u << digit_bits;
    const uint32_t words = n_passes * radix;
    for (uint32_t k = threadIdx.x; k < words; k += blockDim.x) s_hist[k] = 0u;
    if (threadIdx.x < n_passes) s_shift[threadIdx.x] = shifts_dev[threadIdx.x];
    __syncthreads();
    const uint32_t n = *n_dev;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t key = keys[i];
        for (uint32_t p = 0; p < n_passes; p++)
            atomicAdd(&s_hist[p * radix + ((uint32_t)(key >> s_shift[p]) & (radix - 1u))], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < words; k += blockDim.x)
        if (s_hist[k]) atomicAdd(&hist[k], s_hist[k]);
}

}  // namespace

hipError_t launch_emit(const EmitArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    if (a.n_passes > kMaxPasses || (a.digit_bits != 4 && a.digit_bits != 8)) return hipErrorInvalidValue;
    const uint32_t grid = (a.n + kEmitThreads - 1) / kEmitThreads;
    hipLaunchKernelGGL(k_emit, dim3(grid), dim3(kEmitThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_digit_histogram(const uint64_t* keys, const uint32_t* n_dev, uint32_t n_max,
                                  uint32_t* hist, uint32_t digit_bits, uint32_t n_passes,
                                  const uint32_t* shifts, hipStream_t s) {
    if (n_max == 0) return hipSuccess;
    if (n_passes > kMaxPasses || (digit_bits != 4 && digit_bits != 8)) return hipErrorInvalidValue;
    uint32_t grid = (n_max + 255u) / 256u;
    if (grid > 2048u) grid = 2048u;
    hipLaunchKernelGGL(k_digit_histogram, dim3(grid), dim3(256), 0, s, keys, n_dev, hist, digit_bits,
                       n_passes, shifts);
    return hipGetLastError();
}

}  // namespace gs3d
