// k_composite.hip — the display step after the rasterizer.
//
// k_composite<DEBUG> does the fullscreen composite of the premultiplied splat
// target onto an 8-bit framebuffer: composite_splats.wgsl:25-29 loads the
// texel 1:1, and the pipeline's blend state (GaussianSplatRenderer.ts:939-945)
// is premultiplied "over" on all four channels: out = src + dst * (1 - src.a).
// With DEBUG it first applies the splats-per-tile heat map of
// debug_tile_overlay.wgsl:35-58: a tint from green (few) to red (debug_ref or
// more) at opacity 0.4 over the target, written to the rgba16float debug
// target (the reference's overlay pass output), which is then composited.
//
// Fused here, the heat map and the composite are one pass over the pixels.
// Each thread takes 4 consecutive pixels: one 32-B load of the f16 target,
// one 16-B load and one 16-B store of the framebuffer (+ 32 B of debug target).
// The kernel is HBM-bound: 16 B per pixel (24 B with the heat map).
//
// FMA contraction is off so the results equal the CPU restatement's
// operation by operation (each product and sum rounded to f32).
#include <hip/hip_fp16.h>

#include "gs3d_internal.h"

#pragma clang fp contract(off)

namespace gs3d {
namespace {

constexpr int kCompositeThreads = 256;
constexpr int kCompositePixels = 4;  // per thread

__device__ __forceinline__ float half_bits_to_f32(uint32_t h) {
    return __half2float(__ushort_as_half((unsigned short)h));
}

__device__ __forceinline__ uint32_t f32_to_half_bits(float f) {
    return __half_as_ushort(__float2half_rn(f));
}

// Unorm8 store: clamp to [0, 1], scale by 255, round to nearest.
__device__ __forceinline__ uint32_t to_unorm8(float v) {
    v = fminf(fmaxf(v, 0.0f), 1.0f);
    return (uint32_t)(v * 255.0f + 0.5f);
}

struct CompositeArgs {
    const uint16_t* target;  // W*H x 4 f16 premultiplied (C, 1 - T)
    uint16_t* debug_target;  // W*H x 4 f16 (DEBUG only)
    uint32_t* framebuffer;   // W*H packed unorm8 (byte 0 = R, or B when bgra)
    const uint32_t* tile_offsets;
    uint32_t width, tile_x, tile_y, tiles_x, tiles_y;
    uint32_t pixels;
    float ref;               // max(1, f32(debug_ref))
    uint32_t clear;          // packed unorm8 clear colour in framebuffer byte order
    int use_clear;
    int bgra;
};

template <bool DEBUG>
__device__ __forceinline__ void composite_pixel(const CompositeArgs& a, uint32_t p, uint32_t& h01,
                                                uint32_t& h23, uint32_t& fb) {
    float r = half_bits_to_f32(h01 & 0xFFFFu), g = half_bits_to_f32(h01 >> 16);
    float b = half_bits_to_f32(h23 & 0xFFFFu), al = half_bits_to_f32(h23 >> 16);
    if (DEBUG) {
        const uint32_t y = p / a.width, x = p - y * a.width;
        const uint32_t tx = min(x / a.tile_x, a.tiles_x - 1u), ty = min(y / a.tile_y, a.tiles_y - 1u);
        const uint32_t tile = ty * a.tiles_x + tx;
        const uint32_t lo = a.tile_offsets[tile], hi = a.tile_offsets[tile + 1];
        const uint32_t count = hi >= lo ? hi - lo : 0u;
        const float t = fminf(fmaxf((float)count / a.ref, 0.0f), 1.0f);
        const float tr = fminf(fmaxf(t * 2.0f, 0.0f), 1.0f);
        const float tg = fminf(fmaxf(2.0f - t * 2.0f, 0.0f), 1.0f);
        const float ao = 0.4f, keep = 1.0f - ao;
        const uint32_t o0 = f32_to_half_bits(tr * ao + r * keep);
        const uint32_t o1 = f32_to_half_bits(tg * ao + g * keep);
        const uint32_t o2 = f32_to_half_bits(0.0f * ao + b * keep);
        const uint32_t o3 = f32_to_half_bits(ao + al * keep);
        h01 = o0 | (o1 << 16);
        h23 = o2 | (o3 << 16);
        r = half_bits_to_f32(o0);
        g = half_bits_to_f32(o1);
        b = half_bits_to_f32(o2);
        al = half_bits_to_f32(o3);
    }
    const uint32_t d = a.use_clear ? a.clear : fb;
    // framebuffer bytes -> (R, G, B, A)
    const uint32_t c0 = d & 0xFFu, c1 = (d >> 8) & 0xFFu, c2 = (d >> 16) & 0xFFu, c3 = d >> 24;
    const float dr = (float)(a.bgra ? c2 : c0) / 255.0f;
    const float dg = (float)c1 / 255.0f;
    const float db = (float)(a.bgra ? c0 : c2) / 255.0f;
    const float da = (float)c3 / 255.0f;
    const float inv_a = 1.0f - al;
    const uint32_t orr = to_unorm8(r + dr * inv_a), og = to_unorm8(g + dg * inv_a);
    const uint32_t ob = to_unorm8(b + db * inv_a), oa = to_unorm8(al + da * inv_a);
    fb = (a.bgra ? ob : orr) | (og << 8) | ((a.bgra ? orr : ob) << 16) | (oa << 24);
}

template <bool DEBUG>
__global__ __launch_bounds__(kCompositeThreads) void k_composite(CompositeArgs a) {
    const uint32_t p0 = (blockIdx.x * kCompositeThreads + threadIdx.x) * kCompositePixels;
    if (p0 >= a.pixels) return;
    if (p0 + kCompositePixels <= a.pixels) {
        const uint4* tsrc = reinterpret_cast<const uint4*>(a.target + (size_t)p0 * 4);
        uint4 t0 = tsrc[0], t1 = tsrc[1];
        uint4 fb = a.use_clear ? make_uint4(0, 0, 0, 0) : *reinterpret_cast<const uint4*>(a.framebuffer + p0);
        composite_pixel<DEBUG>(a, p0 + 0, t0.x, t0.y, fb.x);
        composite_pixel<DEBUG>(a, p0 + 1, t0.z, t0.w, fb.y);
        composite_pixel<DEBUG>(a, p0 + 2, t1.x, t1.y, fb.z);
        composite_pixel<DEBUG>(a, p0 + 3, t1.z, t1.w, fb.w);
        if (DEBUG) {
            uint4* dst = reinterpret_cast<uint4*>(a.debug_target + (size_t)p0 * 4);
            dst[0] = t0;
            dst[1] = t1;
        }
        *reinterpret_cast<uint4*>(a.framebuffer + p0) = fb;
        return;
    }
    for (uint32_t p = p0; p < a.pixels; p++) {
        const uint2 t = *reinterpret_cast<const uint2*>(a.target + (size_t)p * 4);
        uint32_t h01 = t.x, h23 = t.y;
        uint32_t fb = a.use_clear ? 0u : a.framebuffer[p];
        composite_pixel<DEBUG>(a, p, h01, h23, fb);
        if (DEBUG) *reinterpret_cast<uint2*>(a.debug_target + (size_t)p * 4) = make_uint2(h01, h23);
        a.framebuffer[p] = fb;
    }
}

}  // namespace

uint32_t pack_unorm8_clear(const float rgba[4], int bgra) {
    uint32_t c[4];
    for (int i = 0; i < 4; i++) {
        float v = rgba[i];
        v = v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f;  // NaN -> 0
        c[i] = (uint32_t)(v * 255.0f + 0.5f);
    }
    return (bgra ? c[2] : c[0]) | (c[1] << 8) | ((bgra ? c[0] : c[2]) << 16) | (c[3] << 24);
}

hipError_t launch_composite(const CompositeParams& p, hipStream_t s) {
    if (p.width == 0 || p.height == 0) return hipSuccess;
    CompositeArgs a{};
    a.target = p.target;
    a.debug_target = p.debug_target;
    a.framebuffer = p.framebuffer;
    a.tile_offsets = p.tile_offsets;
    a.width = p.width;
    a.tile_x = p.tile_x;
    a.tile_y = p.tile_y;
    a.tiles_x = p.tiles_x;
    a.tiles_y = p.tiles_y;
    a.pixels = p.width * p.height;
    a.ref = p.debug_ref > 1u ? (float)p.debug_ref : 1.0f;
    a.clear = p.clear;
    a.use_clear = p.use_clear;
    a.bgra = p.bgra;
    const uint32_t blocks = (uint32_t)composite_launch_shape(a.pixels).grid;
    if (p.debug)
        hipLaunchKernelGGL(k_composite<true>, dim3(blocks), dim3(kCompositeThreads), 0, s, a);
    else
        hipLaunchKernelGGL(k_composite<false>, dim3(blocks), dim3(kCompositeThreads), 0, s, a);
    return hipGetLastError();
}

LaunchShape composite_launch_shape(uint32_t pixels) {
    const uint32_t per_block = kCompositeThreads * kCompositePixels;
    return {(uint32_t)kCompositeThreads, ((uint64_t)pixels + per_block - 1) / per_block};
}

}  // namespace gs3d
