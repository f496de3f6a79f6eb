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
// splat (coalesced id load + three 16-B loads), splats that cannot reach the
// wavefront's pixel rows are dropped by a ballot, and the survivors are
// compacted into the wavefront's LDS slots and broadcast from there. No
// workgroup barriers, and a wavefront leaves as soon as all its pixels are
// saturated. Splats that cannot reach alpha >= 1/255 inside the wavefront's
// rows (circle and opacity-aware ellipse bounds, see may_touch) are skipped
// before the broadcast; this drops no blend the per-pixel rules would keep.
// exp is the hardware exp2 (v_exp_f32) of x*log2(e); results stay within one
// fp16 step of the fp32 restatement on the CPU.
#include <hip/hip_fp16.h>

#include "gs3d_internal.h"

namespace gs3d {

namespace {

__global__ __launch_bounds__(256) void k_identify(const uint64_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ n_dev,
                                                  uint32_t* __restrict__ tile_offsets,
                                                  uint32_t tile_count, uint32_t n_max,
                                                  uint32_t* __restrict__ k_host) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t > tile_count) return;
    uint32_t n = *n_dev;
    if (t == 0 && k_host) *k_host = n;  // K for the host, no copy packet needed
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

// Conservative test of one gathered splat against a wavefront's rectangle of
// pixel centres [rx0, rx1] x [ry0, ry1]. A pixel can only be blended when it is
// inside the radius circle AND alpha = o e^{-q/2} >= 1/255, i.e. inside the
// ellipse q = d^T conic d <= 2 ln(255 o); that ellipse's bounding box has
// half-extents sqrt(t Sigma_xx), sqrt(t Sigma_yy) with Sigma = conic^-1. Both
// shapes are widened by far more than float rounding (1% + 0.05 px), so the
// test only drops splats that cannot pass the exact per-pixel rules.
__device__ __forceinline__ bool may_touch(float4 f0, float4 f1, float rx0, float rx1, float ry0,
                                          float ry1) {
    const float mx = f0.x, my = f0.y, rad = f0.w;
    const float cx = f1.x, cy = f1.y, cz = f1.z, op = f1.w;
    if (!(op * 255.0f >= 0.99f)) return false;  // alpha <= o < 1/255 everywhere
    const float qx = fminf(fmaxf(mx, rx0), rx1);
    const float qy = fminf(fmaxf(my, ry0), ry1);
    const float ex = fabsf(mx - qx), ey = fabsf(my - qy);
    const float rr = rad * 1.01f + 0.05f;
    if (ex * ex + ey * ey > rr * rr) return false;
    const float det = cx * cz - cy * cy;
    const float t = 2.0f * __logf(255.0f * op) * 1.02f + 0.05f;
    if (!(det > 0.0f)) return true;  // degenerate conic: keep (exact rules decide)
    const float hx = sqrtf(t * cz / det) * 1.01f + 0.05f;
    const float hy = sqrtf(t * cx / det) * 1.01f + 0.05f;
    return ex <= hx && ey <= hy;
}

// TW x TH tile; each lane owns PPL pixels of one column, so a wavefront covers
// TW columns x (64 PPL / TW) rows and a workgroup has TW TH / PPL threads.
//
// Per batch of 64 list entries a wavefront gathers one splat per lane, keeps
// the ones that may touch its rows (may_touch) and compacts them, in list
// order, into its own LDS slots as {mean.x, mean.y, r^2, conic.x},
// {2 conic.y, conic.z, opacity, red}, {green, blue}. The composite loop then
// reads slot j with uniform-address LDS loads (a broadcast) and blends
// branch-free: a splat that fails a rule is blended with alpha 0, which leaves
// C and T bit-identical (T * (1 - 0) = T, C + T * 0 * c = C).
template <int TW, int TH, int PPL>
__global__ __launch_bounds__(TW * TH / PPL) void k_rasterize(const float4* __restrict__ proj,
                                                             const uint32_t* __restrict__ values,
                                                             const uint32_t* __restrict__ tile_offsets,
                                                             uint32_t tiles_x, uint32_t width,
                                                             uint32_t height, uint2* __restrict__ image) {
    static_assert(kWave % TW == 0 && (TW * TH) % (kWave * PPL) == 0, "wavefronts cover whole rows");
    constexpr int kLaneRows = kWave / TW;         // rows covered by one pixel slot
    constexpr int kWaveRows = kLaneRows * PPL;    // rows covered by one wavefront
    constexpr int kWaves = TW * TH / (PPL * kWave);
    __shared__ float4 s_a[kWaves][kWave];
    __shared__ float4 s_b[kWaves][kWave];
    __shared__ float2 s_c[kWaves][kWave];

    const uint32_t tile = blockIdx.x;
    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const uint32_t px = tx * TW + (lane % TW);
    const uint32_t py0 = ty * TH + wave * kWaveRows + lane / TW;
    const float pxf = (float)px + 0.5f;

    // pixel-centre rectangle of this wavefront's rows
    const float rx0 = (float)(tx * TW) + 0.5f, rx1 = rx0 + (float)(TW - 1);
    const float ry0 = (float)(ty * TH + wave * kWaveRows) + 0.5f;
    const float ry1 = ry0 + (float)(kWaveRows - 1);

    float c0[PPL], c1[PPL], c2[PPL], tr[PPL], pyf[PPL];
    bool done[PPL];
#pragma unroll
    for (int k = 0; k < PPL; k++) {
        const uint32_t py = py0 + k * kLaneRows;
        c0[k] = c1[k] = c2[k] = 0.0f;
        tr[k] = 1.0f;
        pyf[k] = (float)py + 0.5f;
        done[k] = !(px < width && py < height);
    }
    auto all_done = [&]() {
        bool d = true;
#pragma unroll
        for (int k = 0; k < PPL; k++) d = d && done[k];
        return __ballot(!d) == 0ull;
    };
    const uint32_t start = tile_offsets[tile];
    const uint32_t end = tile_offsets[tile + 1];
    const float min_alpha = 1.0f / 255.0f;

    // software pipeline: batch b+1 is gathered while batch b is composited
    float4 n0 = make_float4(0.f, 0.f, 0.f, 0.f), n1 = n0, n2 = n0;
    if (start < end && start + lane < end) {
        const uint32_t sid = values[start + lane];
        n0 = proj[(size_t)sid * 4 + 0];
        n1 = proj[(size_t)sid * 4 + 1];
        n2 = proj[(size_t)sid * 4 + 2];
    }
    for (uint32_t base = start; base < end; base += kWave) {
        if (all_done()) break;
        const float4 f0 = n0, f1 = n1, f2 = n2;
        const bool have = base + lane < end;
        const uint32_t nidx = base + kWave + lane;
        if (nidx < end) {
            const uint32_t sid = values[nidx];
            n0 = proj[(size_t)sid * 4 + 0];
            n1 = proj[(size_t)sid * 4 + 1];
            n2 = proj[(size_t)sid * 4 + 2];
        }
        const bool hit = have && may_touch(f0, f1, rx0, rx1, ry0, ry1);
        const uint64_t mask = __ballot(hit);
        const int nhit = __popcll(mask);
        if (hit) {
            const int slot = (int)__popcll(mask & ((1ull << lane) - 1ull));
            s_a[wave][slot] = make_float4(f0.x, f0.y, f0.w * f0.w, f1.x);
            s_b[wave][slot] = make_float4(2.0f * f1.y, f1.z, f1.w, f2.x);
            s_c[wave][slot] = make_float2(f2.y, f2.z);
        }
        __builtin_amdgcn_wave_barrier();
        // slot j+1 is read while slot j is blended
        float4 na = s_a[wave][0], nb = s_b[wave][0];
        float2 nc = s_c[wave][0];
        for (int j = 0; j < nhit; j++) {
            const float4 a = na, b = nb;
            const float2 cc = nc;
            if (j + 1 < nhit) {
                na = s_a[wave][j + 1];
                nb = s_b[wave][j + 1];
                nc = s_c[wave][j + 1];
            }
            const float dx = pxf - a.x;
#pragma unroll
            for (int k = 0; k < PPL; k++) {
                const float dy = pyf[k] - a.y;
                const float d2 = dx * dx + dy * dy;
                const float power = -0.5f * (a.w * dx * dx + b.x * dx * dy + b.y * dy * dy);
                float alpha = b.z * __expf(power);
                alpha = alpha > 0.99f ? 0.99f : alpha;
                const bool ok = !done[k] && !(d2 > a.z) && !(power > 0.0f) && !(alpha < min_alpha);
                const float ae = ok ? alpha : 0.0f;
                const float w = tr[k] * ae;
                c0[k] = c0[k] + w * b.w;
                c1[k] = c1[k] + w * cc.x;
                c2[k] = c2[k] + w * cc.y;
                tr[k] = tr[k] * (1.0f - ae);
                done[k] = done[k] || (ok && tr[k] < 1.0e-4f);
            }
            if (all_done()) break;
        }
        __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int k = 0; k < PPL; k++) {
        const uint32_t py = py0 + k * kLaneRows;
        if (px < width && py < height) {
            const uint32_t h0 = __half_as_ushort(__float2half_rn(c0[k]));
            const uint32_t h1 = __half_as_ushort(__float2half_rn(c1[k]));
            const uint32_t h2 = __half_as_ushort(__float2half_rn(c2[k]));
            const uint32_t h3 = __half_as_ushort(__float2half_rn(1.0f - tr[k]));
            image[(size_t)py * width + px] = make_uint2(h0 | (h1 << 16), h2 | (h3 << 16));
        }
    }
}

__global__ void k_f16_to_f32(const uint16_t* in, float* out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = __half2float(__ushort_as_half(in[i]));
}

}  // namespace

hipError_t launch_identify_tiles(const uint64_t* keys, const uint32_t* n_dev, uint32_t n_max,
                                 uint32_t* tile_offsets, uint32_t tile_count, uint32_t* k_host,
                                 hipStream_t s) {
    const uint32_t threads = tile_count + 1u;
    hipLaunchKernelGGL(k_identify, dim3((threads + 255u) / 256u), dim3(256), 0, s, keys, n_dev,
                       tile_offsets, tile_count, n_max, k_host);
    return hipGetLastError();
}

template <int TW, int PPL>
static void launch_raster_variant(uint32_t tiles, const float4* p, const uint32_t* values,
                                  const uint32_t* tile_offsets, uint32_t tiles_x, uint32_t width,
                                  uint32_t height, uint2* img, hipStream_t s) {
    hipLaunchKernelGGL((k_rasterize<TW, TW, PPL>), dim3(tiles), dim3(TW * TW / PPL), 0, s, p, values,
                       tile_offsets, tiles_x, width, height, img);
}

bool raster_variant_exists(uint32_t tile, uint32_t ppl) {
    if (tile != 8 && tile != 16 && tile != 32) return false;
    if (ppl != 1 && ppl != 2 && ppl != 4) return false;
    return tile * tile / ppl >= (uint32_t)kWave;
}

hipError_t launch_rasterize(const gs3d_projected_splat* projected, const uint32_t* values,
                            const uint32_t* tile_offsets, uint32_t tile_x, uint32_t tile_y,
                            uint32_t tiles_x, uint32_t tiles_y, uint32_t width, uint32_t height,
                            uint32_t pixels_per_lane, uint16_t* image_rgba16f, hipStream_t s) {
    const uint32_t tiles = tiles_x * tiles_y;
    if (tiles == 0) return hipSuccess;
    if (tile_x != tile_y || !raster_variant_exists(tile_x, pixels_per_lane)) return hipErrorInvalidValue;
    const float4* p = reinterpret_cast<const float4*>(projected);
    uint2* img = reinterpret_cast<uint2*>(image_rgba16f);
    const int key = tile_x * 10 + pixels_per_lane;
    switch (key) {
        case 81: launch_raster_variant<8, 1>(tiles, p, values, tile_offsets, tiles_x, width, height, img, s); break;
        case 161: launch_raster_variant<16, 1>(tiles, p, values, tile_offsets, tiles_x, width, height, img, s); break;
        case 162: launch_raster_variant<16, 2>(tiles, p, values, tile_offsets, tiles_x, width, height, img, s); break;
        case 164: launch_raster_variant<16, 4>(tiles, p, values, tile_offsets, tiles_x, width, height, img, s); break;
        case 321: launch_raster_variant<32, 1>(tiles, p, values, tile_offsets, tiles_x, width, height, img, s); break;
        case 322: launch_raster_variant<32, 2>(tiles, p, values, tile_offsets, tiles_x, width, height, img, s); break;
        case 324: launch_raster_variant<32, 4>(tiles, p, values, tile_offsets, tiles_x, width, height, img, s); break;
        default: return hipErrorInvalidValue;
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
