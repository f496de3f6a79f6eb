// k_scan.hip — single-pass exclusive scan of per-splat tile counts.
//
// Replaces the three reference passes prefix_scan_local.wgsl:26-47 (Blelloch
// per 2S chunk), prefix_scan_blocks.wgsl:28-43 (one thread scanning chunk sums;
// writes ref_counter[0] and the sentinel offsets[N]) and prefix_scan_add.wgsl
// :22-37 by one decoupled-lookback launch: each workgroup takes a ticket (so
// tiles are processed in launch order), scans its 8192 counts, publishes its
// aggregate, looks back over its predecessors' status words a wavefront (64
// predecessors) at a time, then writes final offsets. The result is the same
// exclusive sum mod 2^32 with offsets[N] = ref_counter[0] = total.
// With `order` it scans counts[order[i]] (the counts in depth order).
// HBM traffic: 4 B in + 4 B out per splat (+ 4 B of order).
#include "gs3d_internal.h"

namespace gs3d {

namespace {

__device__ __forceinline__ uint64_t load_status(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Inclusive wave scan (64 lanes) of u32.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t t = __shfl_up(v, d, kWave);
        if (lane >= d) v += t;
    }
    return v;
}

// The in-range paths of both scans load (and k_scan stores) a thread's items
// as uint4.
static_assert(kScanItems % 4 == 0, "k_scan and k_scan_rects load a thread's items as uint4");

__global__ __launch_bounds__(kScanThreads) void k_scan(const uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ order,
                                                       uint32_t* __restrict__ offsets,
                                                       uint32_t* __restrict__ ref_counter,
                                                       uint64_t* status, uint32_t* ticket,
                                                       uint32_t n, uint32_t epoch) {
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_wave_sum[kScanThreads / kWave];
    __shared__ uint32_t s_prefix;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    if (tid == 0) s_part = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t part = s_part;
    const size_t base = (size_t)part * kScanTile + (size_t)tid * kScanItems;

    uint32_t v[kScanItems];
    if (order) {
#pragma unroll
        for (int k = 0; k < kScanItems; k++) v[k] = (base + k < n) ? counts[order[base + k]] : 0u;
    } else if (base + kScanItems <= n) {
        const uint4* src = reinterpret_cast<const uint4*>(counts + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; q++) {
            const uint4 x = src[q];
            v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; k++) v[k] = (base + k < n) ? counts[base + k] : 0u;
    }
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) local += v[k];
    const uint32_t incl = wave_inclusive_scan(local);
    if (lane == kWave - 1) s_wave_sum[wave] = incl;
    __syncthreads();
    uint32_t wave_prefix = 0, tile_total = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / kWave; w++) {
        const uint32_t ws = s_wave_sum[w];
        if (w < wave) wave_prefix += ws;
        tile_total += ws;
    }

    if (wave == 0) {
        // Publish, then look back 64 predecessors at a time.
        if (lane == 0) {
            store_status(&status[part],
                         status_word(epoch, part == 0 ? kFlagPrefix : kFlagAggregate, tile_total));
        }
        uint32_t excl = 0;
        int64_t p = (int64_t)part - 1;
        while (p >= 0) {
            const int64_t q = p - lane;
            uint64_t w = 0;
            uint32_t flag = 0;
            if (q >= 0) {
                do {
                    w = load_status(&status[q]);
                    flag = ((uint32_t)(w >> 34) == epoch) ? (uint32_t)((w >> 32) & 3u) : 0u;
                } while (flag == 0);
            }
            // lanes with q < 0 contribute nothing and act as a prefix stop.
            const uint64_t stop = __ballot(q < 0 || flag == kFlagPrefix);
            const int first = stop ? __ffsll((unsigned long long)stop) - 1 : kWave;
            uint32_t val = (q >= 0 && lane <= first) ? (uint32_t)w : 0u;
#pragma unroll
            for (int d = kWave / 2; d > 0; d >>= 1) val += __shfl_xor(val, d, kWave);
            excl += val;
            if (stop) break;
            p -= kWave;
        }
        if (lane == 0) {
            if (part != 0) store_status(&status[part], status_word(epoch, kFlagPrefix, excl + tile_total));
            s_prefix = excl;
            const uint32_t parts = (n + kScanTile - 1) / kScanTile;
            if (part == parts - 1 || n == 0) {
                offsets[n] = excl + tile_total;
                ref_counter[0] = excl + tile_total;
            }
        }
    }
    __syncthreads();
    uint32_t run = s_prefix + wave_prefix + (incl - local);
    uint32_t o[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        o[k] = run;
        run += v[k];
    }
    if (base + kScanItems <= n) {
        uint4* dst = reinterpret_cast<uint4*>(offsets + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; q++) dst[q] = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; k++)
            if (base + k < n) offsets[base + k] = o[k];
    }
}

// The frame's scan: the counts are not read from memory but counted from the
// tile rectangles of the splats in depth order: gathered through `order` (8 B
// per visible splat), or, with rects8, read in order (4 B packed rectangles
// that the last depth pass carried into depth order). Culled splats have
// depth key 0xFFFFFFFF and count 0.
// Besides the offsets it writes what emit needs per position, so emit reads
// records in order instead of gathering splats: recs[p] = (first key, first
// tile id, width in tiles, depth key); recs[n].x = total = ref_counter[0].
__global__ __launch_bounds__(kScanThreads) void k_scan_rects(ScanRectsArgs a) {
    const uint32_t epoch = resolve_epoch(a.epoch, a.epoch_base);
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_wave_sum[kScanThreads / kWave];
    __shared__ uint32_t s_prefix;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const uint32_t n = a.n;
    if (tid == 0) s_part = atomicAdd(a.ticket, 1u);
    __syncthreads();
    const uint32_t part = s_part;
    const uint32_t parts = n == 0 ? 1u : (n + kScanTile - 1) / kScanTile;
    if (part >= parts) return;
    const size_t base = (size_t)part * kScanTile + (size_t)tid * kScanItems;

    // per position: the splat id (the gather path) or the packed rectangle (rects8)
    uint32_t sid[kScanItems], dk[kScanItems];
    const uint32_t* first = a.rects8 ? a.rects8 : a.order;
    if (base + kScanItems <= n) {
        const uint4* o = reinterpret_cast<const uint4*>(first + base);
        const uint4* d = reinterpret_cast<const uint4*>(a.depth_keys + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; q++) {
            const uint4 oq = o[q], dq = d[q];
            sid[4 * q] = oq.x; sid[4 * q + 1] = oq.y; sid[4 * q + 2] = oq.z; sid[4 * q + 3] = oq.w;
            dk[4 * q] = dq.x; dk[4 * q + 1] = dq.y; dk[4 * q + 2] = dq.z; dk[4 * q + 3] = dq.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; k++) {
            sid[k] = base + k < n ? first[base + k] : 0u;
            dk[k] = base + k < n ? a.depth_keys[base + k] : 0xFFFFFFFFu;
        }
    }
    uint32_t v[kScanItems], t0[kScanItems], wd[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        v[k] = 0u;
        t0[k] = 0u;
        wd[k] = 0u;
        if (dk[k] != 0xFFFFFFFFu) {
            uint32_t x0, y0, x1, y1;
            if (a.rects8) {
                const uint32_t r = sid[k];
                x0 = r & 0xFFu, y0 = (r >> 8) & 0xFFu, x1 = (r >> 16) & 0xFFu, y1 = r >> 24;
            } else {
                const uint2 r = a.rects[sid[k]];
                x0 = r.x & 0xFFFFu, y0 = r.x >> 16, x1 = r.y & 0xFFFFu, y1 = r.y >> 16;
            }
            wd[k] = x1 - x0 + 1u;
            v[k] = wd[k] * (y1 - y0 + 1u);
            t0[k] = y0 * a.tiles_x + x0;
        }
    }
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) local += v[k];
    const uint32_t incl = wave_inclusive_scan(local);
    if (lane == kWave - 1) s_wave_sum[wave] = incl;
    __syncthreads();
    uint32_t wave_prefix = 0, tile_total = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / kWave; w++) {
        const uint32_t ws = s_wave_sum[w];
        if (w < wave) wave_prefix += ws;
        tile_total += ws;
    }

    if (wave == 0) {
        // Publish, then look back 64 predecessors at a time.
        if (lane == 0) {
            store_status(&a.status[part],
                         status_word(epoch, part == 0 ? kFlagPrefix : kFlagAggregate, tile_total));
        }
        uint32_t excl = 0;
        int64_t p = (int64_t)part - 1;
        while (p >= 0) {
            const int64_t q = p - lane;
            uint64_t w = 0;
            uint32_t flag = 0;
            if (q >= 0) {
                do {
                    w = load_status(&a.status[q]);
                    flag = ((uint32_t)(w >> 34) == epoch) ? (uint32_t)((w >> 32) & 3u) : 0u;
                } while (flag == 0);
            }
            // lanes with q < 0 contribute nothing and act as a prefix stop.
            const uint64_t stop = __ballot(q < 0 || flag == kFlagPrefix);
            const int first = stop ? __ffsll((unsigned long long)stop) - 1 : kWave;
            uint32_t val = (q >= 0 && lane <= first) ? (uint32_t)w : 0u;
#pragma unroll
            for (int d = kWave / 2; d > 0; d >>= 1) val += __shfl_xor(val, d, kWave);
            excl += val;
            if (stop) break;
            p -= kWave;
        }
        if (lane == 0) {
            if (part != 0) store_status(&a.status[part], status_word(epoch, kFlagPrefix, excl + tile_total));
            s_prefix = excl;
            if (part == parts - 1) {
                a.recs[n] = make_uint4(excl + tile_total, 0u, 0u, 0u);
                a.ref_counter[0] = excl + tile_total;
            }
        }
    }
    __syncthreads();
    uint32_t run = s_prefix + wave_prefix + (incl - local);
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) a.recs[base + k] = make_uint4(run, t0[k], wd[k], dk[k]);
        run += v[k];
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t value, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = value;
}

}  // namespace

hipError_t launch_scan(const uint32_t* counts, const uint32_t* order, uint32_t* offsets,
                       uint32_t* ref_counter, uint64_t* status, uint32_t* ticket, uint32_t n,
                       uint32_t epoch, hipStream_t s) {
    const uint32_t parts = n == 0 ? 1u : (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan, dim3(parts), dim3(kScanThreads), 0, s, counts, order, offsets,
                       ref_counter, status, ticket, n, epoch);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, s, p, value, n);
    return hipGetLastError();
}

LaunchShape scan_rects_launch_shape(uint32_t n) {
    return {(uint32_t)kScanThreads, n == 0 ? 1ull : ((uint64_t)n + kScanTile - 1) / kScanTile};
}

hipError_t launch_scan_rects(const ScanRectsArgs& a, hipStream_t s) {
    const uint32_t parts = (uint32_t)scan_rects_launch_shape(a.n).grid;
    hipLaunchKernelGGL(k_scan_rects, dim3(parts), dim3(kScanThreads), 0, s, a);
    return hipGetLastError();
}

}  // namespace gs3d
