// k_preprocess.hip — per-splat EWA projection, SH colour and tile-overlap count.
//
// Follows preprocess_splats.wgsl:75-184 (reference repo, src/shaders/compute/):
// clip/NDC centre cull, screen mapping, cov3d = (R S)(R S)^T, J W Sigma W^T J^T
// without a frustum clamp, +0.3 low-pass on the diagonal, det <= 0 cull,
// conic = (c, -b, a)/det, radius = max(2, ceil(3 sqrt(lambda_max))), tile rect
// clamped to the grid, count = w*h. Colour: degree 0 takes the loader's
// activated colour (the reference behaviour); degree 1-3 evaluates the real SH
// basis at normalize(pos - camera.position) and clamps 0.5 + sum to [0, 1].
//
// One thread per splat; HBM-bound: 64 B of vec4 input (+ 4*(deg+1)^2*3 B of
// SH, coefficient-major SoA so every load is coalesced) in, 64 + 4 B out.
// FMA contraction is off in this file so the float results equal the CPU
// restatement operation for operation (same association as the WGSL).
#include <hip/hip_fp16.h>

#include <algorithm>

#include "gs3d_internal.h"

#pragma clang fp contract(off)

namespace gs3d {

namespace {

constexpr float kC0 = 0.28209479177387814f;
constexpr float kC1 = 0.4886025119029199f;
constexpr float kC2_0 = 1.0925484305920792f;
constexpr float kC2_1 = -1.0925484305920792f;
constexpr float kC2_2 = 0.31539156525252005f;
constexpr float kC2_3 = -1.0925484305920792f;
constexpr float kC2_4 = 0.5462742152960396f;
constexpr float kC3_0 = -0.5900435899266435f;
constexpr float kC3_1 = 2.890611442640554f;
constexpr float kC3_2 = -0.4570457994644658f;
constexpr float kC3_3 = 0.3731763325901154f;
constexpr float kC3_4 = -0.4570457994644658f;
constexpr float kC3_5 = 1.445305721320277f;
constexpr float kC3_6 = -0.5900435899266435f;

// Basis values b[0..15] at unit direction (x, y, z) for degree DEG.
template <int DEG>
__device__ __forceinline__ void sh_basis(float x, float y, float z, float* b) {
    b[0] = kC0;
    if (DEG > 0) {
        b[1] = -kC1 * y;
        b[2] = kC1 * z;
        b[3] = -kC1 * x;
    }
    if (DEG > 1) {
        float xx = x * x, yy = y * y, zz = z * z;
        float xy = x * y, yz = y * z, xz = x * z;
        b[4] = kC2_0 * xy;
        b[5] = kC2_1 * yz;
        b[6] = kC2_2 * (2.0f * zz - xx - yy);
        b[7] = kC2_3 * xz;
        b[8] = kC2_4 * (xx - yy);
        if (DEG > 2) {
            b[9] = kC3_0 * y * (3.0f * xx - yy);
            b[10] = kC3_1 * xy * z;
            b[11] = kC3_2 * y * (4.0f * zz - xx - yy);
            b[12] = kC3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
            b[13] = kC3_4 * x * (4.0f * zz - xx - yy);
            b[14] = kC3_5 * z * (xx - yy);
            b[15] = kC3_6 * x * (xx - 3.0f * yy);
        }
    }
}

// Colour of one splat seen from camera position `eye`: 0.5 + sum_k b_k c_k per
// channel, clamped to [0, 1]. src.coef(k, ch) gives SH coefficient k of channel ch.
template <int DEG, class Src>
__device__ __forceinline__ void sh_color(const float* eye, float4 p, const Src& src, float* rgb) {
    float dx = p.x - eye[0];
    float dy = p.y - eye[1];
    float dz = p.z - eye[2];
    float len = sqrtf(dx * dx + dy * dy + dz * dz);
    float inv = 1.0f / len;
    float b[16];
    sh_basis<DEG>(dx * inv, dy * inv, dz * inv, b);
    constexpr int kCoef = (DEG + 1) * (DEG + 1);
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        float acc = b[0] * src.coef(0, ch);
#pragma unroll
        for (int k = 1; k < kCoef; k++) acc = acc + b[k] * src.coef(k, ch);
        float v = acc + 0.5f;
        rgb[ch] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    }
}

// Splat attributes read from global memory when first needed (culled splats
// never load their rotation, scale, colour or SH).
struct GlobalSplat {
    const PreprocessArgs& a;
    uint32_t i;
    __device__ float4 rot() const { return a.rot[i]; }
    __device__ float4 scale() const { return a.scale[i]; }
    __device__ float4 color() const { return a.color[i]; }
    __device__ float coef(int k, int ch) const { return a.sh[(size_t)(k * 3 + ch) * a.n + i]; }
};

// Splat attributes already in registers (one splat projected into several views).
template <int DEG>
struct RegSplat {
    float4 q, s, c;
    float sh[3][(DEG + 1) * (DEG + 1)];
    __device__ float4 rot() const { return q; }
    __device__ float4 scale() const { return s; }
    __device__ float4 color() const { return c; }
    __device__ float coef(int k, int ch) const { return sh[ch][k]; }
};

struct ViewGeometry {
    const gs3d_camera& cam;
    uint32_t tiles_x, tiles_y;
    float inv_tile_x, inv_tile_y;
};

// preprocess_splats.wgsl:75-184 for one splat in one view. Returns the tile
// count (0 when culled; then o0..o3 are zero).
template <int DEG, class Src>
__device__ __forceinline__ uint32_t project_splat(const ViewGeometry& g, float4 p, const Src& src,
                                                  float4& o0, float4& o1, float4& o2, float4& o3) {
    const float* V = g.cam.view;
    const float* P = g.cam.projection;
    float vp[4], cp[4];
#pragma unroll
    for (int r = 0; r < 4; r++) vp[r] = V[r] * p.x + V[4 + r] * p.y + V[8 + r] * p.z + V[12 + r] * 1.0f;
#pragma unroll
    for (int r = 0; r < 4; r++)
        cp[r] = P[r] * vp[0] + P[4 + r] * vp[1] + P[8 + r] * vp[2] + P[12 + r] * vp[3];

    bool ok = cp[3] > 0.0f;
    float ndc0 = cp[0] / cp[3], ndc1 = cp[1] / cp[3], ndc2 = cp[2] / cp[3];
    ok = ok && !(ndc0 < -1.0f || ndc0 > 1.0f || ndc1 < -1.0f || ndc1 > 1.0f || ndc2 < 0.0f ||
                 ndc2 > 1.0f);
    const float W = g.cam.viewport[0], H = g.cam.viewport[1];
    uint32_t count = 0;
    o0 = make_float4(0.f, 0.f, 0.f, 0.f), o1 = o0, o2 = o0, o3 = o0;
    if (ok) {
        const float sx = (ndc0 * 0.5f + 0.5f) * W;
        const float sy = (1.0f - (ndc1 * 0.5f + 0.5f)) * H;

        const float4 q = src.rot();
        const float4 s = src.scale();
        const float qw = q.x, qx = q.y, qy = q.z, qz = q.w;
        const float xx = qx * qx, yy = qy * qy, zz = qz * qz;
        const float xy = qx * qy, xz = qx * qz, yz = qy * qz;
        const float wx = qw * qx, wy = qw * qy, wz = qw * qz;
        // R[col][row], columns as in quat_to_mat3 (:60-64)
        const float R[3][3] = {
            {1.0f - 2.0f * (yy + zz), 2.0f * (xy + wz), 2.0f * (xz - wy)},
            {2.0f * (xy - wz), 1.0f - 2.0f * (xx + zz), 2.0f * (yz + wx)},
            {2.0f * (xz + wy), 2.0f * (yz - wx), 1.0f - 2.0f * (xx + yy)},
        };
        const float sc[3] = {s.x, s.y, s.z};
        float M[3][3];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int r = 0; r < 3; r++) M[c][r] = R[c][r] * sc[c];
        float S[3][3];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int r = 0; r < 3; r++) S[c][r] = M[0][r] * M[0][c] + M[1][r] * M[1][c] + M[2][r] * M[2][c];

        const float fx = P[0] * W * 0.5f;
        const float fy = P[5] * H * 0.5f;
        const float zc = -vp[2];
        const float iz = 1.0f / zc;
        const float iz2 = iz * iz;
        const float J[3][3] = {
            {fx * iz, 0.0f, 0.0f},
            {0.0f, -fy * iz, 0.0f},
            {fx * vp[0] * iz2, -fy * vp[1] * iz2, 0.0f},
        };
        float T[3][3];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int r = 0; r < 3; r++)
                T[c][r] = J[0][r] * V[c * 4 + 0] + J[1][r] * V[c * 4 + 1] + J[2][r] * V[c * 4 + 2];
        float TS[3][3];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int r = 0; r < 3; r++) TS[c][r] = T[0][r] * S[c][0] + T[1][r] * S[c][1] + T[2][r] * S[c][2];
        const float c00 = TS[0][0] * T[0][0] + TS[1][0] * T[1][0] + TS[2][0] * T[2][0];
        const float c01 = TS[0][1] * T[0][0] + TS[1][1] * T[1][0] + TS[2][1] * T[2][0];
        const float c11 = TS[0][1] * T[0][1] + TS[1][1] * T[1][1] + TS[2][1] * T[2][1];
        const float ca = c00 + 0.3f;
        const float cb = c01;
        const float cc = c11 + 0.3f;
        const float det = ca * cc - cb * cb;
        if (det > 0.0f) {
            const float idet = 1.0f / det;
            const float mid = 0.5f * (ca + cc);
            const float disc = mid * mid - det;
            const float lam = mid + sqrtf(disc > 0.1f ? disc : 0.1f);
            float rad = ceilf(3.0f * sqrtf(lam));
            if (rad < 2.0f) rad = 2.0f;
            float minx = sx - rad, miny = sy - rad;
            minx = minx < 0.0f ? 0.0f : minx;
            miny = miny < 0.0f ? 0.0f : miny;
            float maxx = sx + rad, maxy = sy + rad;
            maxx = maxx > W - 1.0f ? W - 1.0f : maxx;
            maxy = maxy > H - 1.0f ? H - 1.0f : maxy;
            uint32_t t0x = (uint32_t)floorf(minx * g.inv_tile_x);
            uint32_t t0y = (uint32_t)floorf(miny * g.inv_tile_y);
            uint32_t t1x = (uint32_t)floorf(maxx * g.inv_tile_x);
            uint32_t t1y = (uint32_t)floorf(maxy * g.inv_tile_y);
            t0x = min(t0x, g.tiles_x - 1u);
            t0y = min(t0y, g.tiles_y - 1u);
            t1x = min(t1x, g.tiles_x - 1u);
            t1y = min(t1y, g.tiles_y - 1u);
            if (t0x <= t1x && t0y <= t1y) {
                count = (t1x - t0x + 1u) * (t1y - t0y + 1u);
                const float4 col = src.color();
                float rgb[3] = {col.x, col.y, col.z};
                if (DEG > 0) sh_color<DEG>(g.cam.position, p, src, rgb);
                o0 = make_float4(sx, sy, vp[2], rad);
                o1 = make_float4(cc * idet, -cb * idet, ca * idet, col.w);
                o2 = make_float4(rgb[0], rgb[1], rgb[2], __uint_as_float(1u));
                o3 = make_float4(__uint_as_float(t0x), __uint_as_float(t0y), __uint_as_float(t1x),
                                 __uint_as_float(t1y));
            }
        }
    }
    return count;
}

__device__ __forceinline__ void store_projected(gs3d_projected_splat* out_base, uint32_t i, uint32_t count,
                                                const float4& o0, const float4& o1, const float4& o2,
                                                const float4& o3, uint32_t* depth_keys, uint2* rects,
                                                uint32_t* rects8 = nullptr) {
    float4* out = reinterpret_cast<float4*>(out_base + i);
    out[0] = o0;
    out[1] = o1;
    out[2] = o2;
    out[3] = o3;
    if (depth_keys) depth_keys[i] = count ? __float_as_uint(-o0.z) : 0xFFFFFFFFu;
    // tile coordinates are < 2048 (viewport <= 16384 px, tiles >= 8 px)
    if (rects)
        rects[i] = make_uint2(__float_as_uint(o3.x) | (__float_as_uint(o3.y) << 16),
                              __float_as_uint(o3.z) | (__float_as_uint(o3.w) << 16));
    // the host sets rects8 only when every tile coordinate is < 256
    if (rects8)
        rects8[i] = __float_as_uint(o3.x) | (__float_as_uint(o3.y) << 8) | (__float_as_uint(o3.z) << 16) |
                    (__float_as_uint(o3.w) << 24);
}

template <int DEG>
__global__ __launch_bounds__(256) void k_preprocess(PreprocessArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    // the frame's counters (tickets, histograms, K) start from zero; nothing
    // reads them before this kernel ends
    if (a.counters && i < a.counter_words) a.counters[i] = (i == a.count_slot) ? a.n : 0u;
    if (i >= a.n) return;
    const float4 p = a.pos[i];
    const ViewGeometry g{a.cam, a.tiles_x, a.tiles_y, a.inv_tile_x, a.inv_tile_y};
    float4 o0, o1, o2, o3;
    const uint32_t count = project_splat<DEG>(g, p, GlobalSplat{a, i}, o0, o1, o2, o3);
    store_projected(a.out, i, count, o0, o1, o2, o3, a.depth_keys, a.rects, a.rects8);
    a.counts[i] = count;
}

// Views one workgroup of the multi-view preprocess projects each splat into:
// the splat's attributes and SH coefficients are read once per kViewsPerBlock
// views (measured on 64 views of 2M SH3 splats: 2,803 / 2,925 / 2,964 / 2,975
// / 3,015 views/s at 4 / 8 / 16 / 32 / 64). The learned-colour kernel
// prefetches one colour per view and keeps 8 (1,722 views/s at 8, 1,700 at 32,
// whose prefetch takes 167 VGPRs).
constexpr uint32_t kViewsPerBlock = 64;
constexpr uint32_t kViewsPerBlockColors = 8;

template <bool VIEW_COLORS>
constexpr uint32_t views_per_block() { return VIEW_COLORS ? kViewsPerBlockColors : kViewsPerBlock; }

// One splat into the views [view0, view0 + views_per_block) of a batch: the
// splat's attributes and SH coefficients are read once, then projected and
// coloured per view with the same arithmetic as k_preprocess. Each view's
// counters are zeroed (N stored) by the first threads of the grid's x range.
//
// VIEW_COLORS (learned colours, degree 0): each view's colour comes from
// view_colors instead of the DC term; the block's views' colours are loaded up
// front, so their latency overlaps instead of stalling every view in turn. A
// separate instantiation, so the plain degree-0 kernel keeps its registers.
//
// A batch view's records are read only by its rasterizer, which reads the
// first 48 B, and only for splats in its tile lists: each visible splat gets a
// 48-B record (centre/depth/radius, conic/opacity, colour) and its rectangle;
// a culled splat only its depth key 0xFFFFFFFF. FULL writes the 64-B
// gs3d_projected_splat of every splat instead, as k_preprocess does.
template <int DEG, bool VIEW_COLORS, bool FULL>
__global__ __launch_bounds__(256) void k_preprocess_views(MultiPreprocessArgs a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    constexpr uint32_t VPB = views_per_block<VIEW_COLORS>();
    const uint32_t v0 = blockIdx.y * VPB;
    const uint32_t v1 = min(v0 + VPB, a.n_views);
    if (i < a.counter_words)
        for (uint32_t v = v0; v < v1; v++) a.views[v].counters[i] = (i == a.count_slot) ? a.n : 0u;
    if (i >= a.n) return;
    const float4 p = a.pos[i];
    RegSplat<DEG> r;
    r.q = a.rot[i];
    r.s = a.scale[i];
    r.c = a.color[i];
    if (DEG > 0) {
#pragma unroll
        for (int k = 0; k < (DEG + 1) * (DEG + 1); k++)
#pragma unroll
            for (int ch = 0; ch < 3; ch++) r.sh[ch][k] = a.sh[(size_t)(k * 3 + ch) * a.n + i];
    }
    // one view; with `has_col`, `col` is the splat's colour for this view (3 halves)
    auto do_view = [&](uint32_t v, bool has_col, uint2 col) {
        const ViewParams& vp = a.views[v];
        const ViewGeometry g{vp.cam, vp.tiles_x, vp.tiles_y, vp.inv_tile_x, vp.inv_tile_y};
        float4 o0, o1, o2, o3;
        const uint32_t count = project_splat<DEG>(g, p, r, o0, o1, o2, o3);
        if (has_col && count) {
            o2.x = __half2float(__ushort_as_half((unsigned short)(col.x & 0xFFFFu)));
            o2.y = __half2float(__ushort_as_half((unsigned short)(col.x >> 16)));
            o2.z = __half2float(__ushort_as_half((unsigned short)(col.y & 0xFFFFu)));
        }
        if constexpr (FULL) {
            store_projected(reinterpret_cast<gs3d_projected_splat*>(vp.out), i, count, o0, o1, o2, o3,
                            vp.depth_keys, vp.rects, vp.rects8);
        } else if (count) {
            float4* out = vp.out + (size_t)i * 3;
            out[0] = o0;
            out[1] = o1;
            out[2] = o2;
            vp.depth_keys[i] = __float_as_uint(-o0.z);
            if (vp.rects)
                vp.rects[i] = make_uint2(__float_as_uint(o3.x) | (__float_as_uint(o3.y) << 16),
                                         __float_as_uint(o3.z) | (__float_as_uint(o3.w) << 16));
            if (vp.rects8)
                vp.rects8[i] = __float_as_uint(o3.x) | (__float_as_uint(o3.y) << 8) |
                               (__float_as_uint(o3.z) << 16) | (__float_as_uint(o3.w) << 24);
        } else {
            vp.depth_keys[i] = 0xFFFFFFFFu;
        }
    };
    if constexpr (VIEW_COLORS) {
        uint2 vc[VPB];
#pragma unroll
        for (int k = 0; k < (int)VPB; k++)
            vc[k] = (v0 + k < v1) ? a.view_colors[(size_t)(v0 + k) * a.n + i] : make_uint2(0u, 0u);
#pragma unroll
        for (int k = 0; k < (int)VPB; k++)
            if (v0 + k < v1) do_view(v0 + k, true, vc[k]);
    } else {
        for (uint32_t v = v0; v < v1; v++) do_view(v, false, make_uint2(0u, 0u));
    }
}

// f_dc (N x 3) + rest (N x 3R, channel-major per splat) -> coefficient-major SoA.
__global__ __launch_bounds__(256) void k_sh_relayout(const float* f_dc, const float* rest, float* sh,
                                                     uint32_t n, int ncoef) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const int R = ncoef - 1;
    for (int ch = 0; ch < 3; ch++) {
        sh[(size_t)ch * n + i] = f_dc[(size_t)i * 3 + ch];
        for (int k = 1; k < ncoef; k++)
            sh[(size_t)(k * 3 + ch) * n + i] = rest[(size_t)i * 3 * R + (size_t)ch * R + (k - 1)];
    }
}

}  // namespace

LaunchShape preprocess_launch_shape(uint32_t n) { return {256u, (n + 255ull) / 256u}; }

hipError_t launch_preprocess(const PreprocessArgs& a, hipStream_t s) {
    const uint32_t threads = a.n > (a.counters ? a.counter_words : 0u) ? a.n : a.counter_words;
    if (threads == 0) return hipSuccess;
    const uint32_t grid = (uint32_t)preprocess_launch_shape(threads).grid;
    switch (a.sh_degree) {
        case 0: hipLaunchKernelGGL(k_preprocess<0>, dim3(grid), dim3(256), 0, s, a); break;
        case 1: hipLaunchKernelGGL(k_preprocess<1>, dim3(grid), dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_preprocess<2>, dim3(grid), dim3(256), 0, s, a); break;
        case 3: hipLaunchKernelGGL(k_preprocess<3>, dim3(grid), dim3(256), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool FULL>
static hipError_t launch_views(const MultiPreprocessArgs& a, hipStream_t s) {
    const uint32_t threads = std::max(a.n, a.counter_words);
    if (threads == 0 || a.n_views == 0) return hipSuccess;
    const uint32_t vpb = views_per_block<false>(), vpb_colors = views_per_block<true>();
    const dim3 grid((threads + 255u) / 256u, (a.n_views + vpb - 1) / vpb);
    // learned colours replace the SH evaluation (the DC colour path is the
    // cheapest projection; its colour is then overwritten)
    if (a.view_colors) {
        const dim3 grid_colors(grid.x, (a.n_views + vpb_colors - 1) / vpb_colors);
        hipLaunchKernelGGL((k_preprocess_views<0, true, FULL>), grid_colors, dim3(256), 0, s, a);
        return hipGetLastError();
    }
    switch (a.sh_degree) {
        case 0: hipLaunchKernelGGL((k_preprocess_views<0, false, FULL>), grid, dim3(256), 0, s, a); break;
        case 1: hipLaunchKernelGGL((k_preprocess_views<1, false, FULL>), grid, dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_preprocess_views<2, false, FULL>), grid, dim3(256), 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_preprocess_views<3, false, FULL>), grid, dim3(256), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_preprocess_views(const MultiPreprocessArgs& a, hipStream_t s) { return launch_views<false>(a, s); }

hipError_t launch_preprocess_views_full(const MultiPreprocessArgs& a, hipStream_t s) {
    return launch_views<true>(a, s);
}

hipError_t launch_sh_relayout(const float* f_dc, const float* rest, float* sh, uint32_t n, int deg,
                              hipStream_t s) {
    if (n == 0) return hipSuccess;
    const int ncoef = (deg + 1) * (deg + 1);
    hipLaunchKernelGGL(k_sh_relayout, dim3((n + 255u) / 256u), dim3(256), 0, s, f_dc, rest, sh, n,
                       ncoef);
    return hipGetLastError();
}

}  // namespace gs3d
