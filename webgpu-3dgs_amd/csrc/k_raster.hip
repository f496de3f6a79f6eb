// k_raster.hip — tile ranges and the tiled front-to-back rasterizer.
//
// identify: the reference (identify_tile_ranges.wgsl:30-52) runs one thread
// per sorted key; at a tile START it fills tile_offsets[prev+1 .. curr] = i
// serially, at a tile END it writes tile_offsets[curr+1] = i+1, and tiles past
// the last populated one keep the 0 of the per-rebin clear. That array equals,
// for every tile t: lower_bound(key.hi >= t) if t <= last+1, else 0 (last =
// tile of the final key; all zero when there are no keys). One thread per tile
// evaluates that closed form with a binary search over the sorted keys, so no
// thread walks a long gap serially.
//
// rasterize: same per-pixel arithmetic and order as rasterize_tiles.wgsl:43-110
// (pixel centre +0.5; skip if |d|^2 > r^2; power = -1/2 d^T conic d, skip if
// power > 0; alpha = min(0.99, o e^power), skip if alpha < 1/255; C += T a c,
// T *= 1-a, THEN stop when T < 1e-4; out = (C, 1-T) as rgba16float). The
// schedule is CDNA-native: each wavefront owns 64 pixels of the tile and walks
// the tile's sorted list on its own, 64 splats per step: one lane gathers one
// splat (coalesced id load + three 16-B loads), splats whose circle misses the
// wavefront's pixel rows are dropped by a ballot, and the survivors are
// broadcast lane by lane through v_readlane into scalar registers. No LDS, no
// workgroup barriers, and a wavefront leaves as soon as all 64 pixels are
// saturated.
#include <hip/hip_fp16.h>

#include "gs3d_internal.h"

#pragma clang fp contract(off)

namespace gs3d {

namespace {

__global__ __launch_bounds__(256) void k_identify(const uint64_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ n_dev,
                                                  uint32_t* __restrict__ tile_offsets,
                                                  uint32_t tile_count, uint32_t n_max) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t > tile_count) return;
    uint32_t n = *n_dev;
    if (n > n_max) n = 0;  // keys did not fit: draw nothing, the host re-runs the frame
    const uint32_t* kw = reinterpret_cast<const uint32_t*>(keys);
    uint32_t out = 0;
    if (n > 0) {
        const uint32_t last = kw[2ull * (n - 1) + 1];
        if (t <= last + 1u) {
            uint32_t lo = 0, hi = n;  // first index with hi-word >= t
            while (lo < hi) {
                const uint32_t mid = lo + ((hi - lo) >> 1);
                if (kw[2ull * mid + 1] < t) lo = mid + 1; else hi = mid;
            }
            out = lo;
        }
    }
    tile_offsets[t] = out;
}

__device__ __forceinline__ float bcast(float x, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}

template <int TW, int TH>
__global__ __launch_bounds__(TW * TH) void k_rasterize(const float4* __restrict__ proj,
                                                       const uint32_t* __restrict__ values,
                                                       const uint32_t* __restrict__ tile_offsets,
                                                       uint32_t tiles_x, uint32_t width,
                                                       uint32_t height, uint2* __restrict__ image) {
    static_assert(TW * TH % kWave == 0 && kWave % TW == 0, "a wavefront covers whole rows");
    constexpr int kRowsPerWave = kWave / TW;
    const uint32_t tile = blockIdx.x;
    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const uint32_t px = tx * TW + (tid % TW);
    const uint32_t py = ty * TH + (tid / TW);
    const bool in_bounds = px < width && py < height;
    const float pxf = (float)px + 0.5f, pyf = (float)py + 0.5f;

    // pixel-centre rectangle of this wavefront's rows
    const float rx0 = (float)(tx * TW) + 0.5f, rx1 = rx0 + (float)(TW - 1);
    const float ry0 = (float)(ty * TH + wave * kRowsPerWave) + 0.5f;
    const float ry1 = ry0 + (float)(kRowsPerWave - 1);

    const uint32_t start = tile_offsets[tile];
    const uint32_t end = tile_offsets[tile + 1];
    float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, tr = 1.0f;
    bool done = !in_bounds;
    const float min_alpha = 1.0f / 255.0f;

    for (uint32_t base = start; base < end; base += kWave) {
        if (__ballot(!done) == 0ull) break;
        const uint32_t idx = base + lane;
        bool hit = idx < end;
        float4 f0 = make_float4(0.f, 0.f, 0.f, 0.f), f1 = f0, f2 = f0;
        if (hit) {
            const uint32_t sid = values[idx];
            f0 = proj[(size_t)sid * 4 + 0];
            f1 = proj[(size_t)sid * 4 + 1];
            f2 = proj[(size_t)sid * 4 + 2];
            // conservative circle-vs-rows test with one pixel of slack
            const float qx = fminf(fmaxf(f0.x, rx0), rx1);
            const float qy = fminf(fmaxf(f0.y, ry0), ry1);
            const float ex = f0.x - qx, ey = f0.y - qy;
            const float rr = f0.w + 1.0f;
            hit = ex * ex + ey * ey <= rr * rr;
        }
        uint64_t mask = __ballot(hit);
        while (mask) {
            const int j = __ffsll((unsigned long long)mask) - 1;
            mask &= mask - 1ull;
            const float mx = bcast(f0.x, j), my = bcast(f0.y, j), rad = bcast(f0.w, j);
            const float cx = bcast(f1.x, j), cy = bcast(f1.y, j), cz = bcast(f1.z, j);
            const float op = bcast(f1.w, j);
            const float r = bcast(f2.x, j), g = bcast(f2.y, j), b = bcast(f2.z, j);
            if (!done) {
                const float dx = pxf - mx, dy = pyf - my;
                if (!(dx * dx + dy * dy > rad * rad)) {
                    const float power = -0.5f * (cx * dx * dx + 2.0f * cy * dx * dy + cz * dy * dy);
                    if (!(power > 0.0f)) {
                        float alpha = op * expf(power);
                        alpha = alpha > 0.99f ? 0.99f : alpha;
                        if (!(alpha < min_alpha)) {
                            const float w = tr * alpha;
                            c0 = c0 + w * r;
                            c1 = c1 + w * g;
                            c2 = c2 + w * b;
                            tr = tr * (1.0f - alpha);
                            if (tr < 1.0e-4f) done = true;
                        }
                    }
                }
            }
            if (__ballot(!done) == 0ull) break;
        }
    }
    if (in_bounds) {
        const uint32_t h0 = __half_as_ushort(__float2half_rn(c0));
        const uint32_t h1 = __half_as_ushort(__float2half_rn(c1));
        const uint32_t h2 = __half_as_ushort(__float2half_rn(c2));
        const uint32_t h3 = __half_as_ushort(__float2half_rn(1.0f - tr));
        image[(size_t)py * width + px] = make_uint2(h0 | (h1 << 16), h2 | (h3 << 16));
    }
}

__global__ void k_f16_to_f32(const uint16_t* in, float* out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = __half2float(__ushort_as_half(in[i]));
}

}  // namespace

hipError_t launch_identify_tiles(const uint64_t* keys, const uint32_t* n_dev, uint32_t n_max,
                                 uint32_t* tile_offsets, uint32_t tile_count, hipStream_t s) {
    const uint32_t threads = tile_count + 1u;
    hipLaunchKernelGGL(k_identify, dim3((threads + 255u) / 256u), dim3(256), 0, s, keys, n_dev,
                       tile_offsets, tile_count, n_max);
    return hipGetLastError();
}

hipError_t launch_rasterize(const gs3d_projected_splat* projected, const uint32_t* values,
                            const uint32_t* tile_offsets, uint32_t tile_x, uint32_t tile_y,
                            uint32_t tiles_x, uint32_t tiles_y, uint32_t width, uint32_t height,
                            uint16_t* image_rgba16f, hipStream_t s) {
    const uint32_t tiles = tiles_x * tiles_y;
    if (tiles == 0) return hipSuccess;
    const float4* p = reinterpret_cast<const float4*>(projected);
    uint2* img = reinterpret_cast<uint2*>(image_rgba16f);
    if (tile_x == 16 && tile_y == 16) {
        hipLaunchKernelGGL((k_rasterize<16, 16>), dim3(tiles), dim3(256), 0, s, p, values, tile_offsets,
                           tiles_x, width, height, img);
    } else if (tile_x == 8 && tile_y == 8) {
        hipLaunchKernelGGL((k_rasterize<8, 8>), dim3(tiles), dim3(64), 0, s, p, values, tile_offsets,
                           tiles_x, width, height, img);
    } else if (tile_x == 32 && tile_y == 32) {
        hipLaunchKernelGGL((k_rasterize<32, 32>), dim3(tiles), dim3(1024), 0, s, p, values,
                           tile_offsets, tiles_x, width, height, img);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_f16_to_f32(const uint16_t* in, float* out, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_f16_to_f32, dim3((uint32_t)blocks), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

}  // namespace gs3d
