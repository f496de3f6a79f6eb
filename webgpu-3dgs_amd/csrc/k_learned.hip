// k_learned.hip — learned view-dependent colour: a per-splat 8-dim latent and a
// shared MLP (latent + unit view direction -> 64 -> 64 -> 3, ReLU, sigmoid) on
// bf16 MFMA with fp32 accumulation, its training step and Adam.
//
// No reference counterpart: the reference colours splats with the DC term
// only (SceneWorker.ts:428-433). The learned colour stands in for the 48
// coefficients of exact SH degree 3, which stays the fallback and the judge.
//
// Matrix layout. Every layer is computed as weights x activations, features
// by samples, with v_mfma_f32_16x16x16bf16_1k on 16x16 tiles:
//   A (16x16 bf16): lane l holds A[l % 16][4 (l / 16) + e], e = 0..3
//   B (16x16 bf16): lane l holds B[4 (l / 16) + e][l % 16]
//   D (16x16 f32):  lane l holds D[4 (l / 16) + e][l % 16]
// so the D tiles of one layer are, element for element, the B tiles of the
// next: the forward pass and the error propagation chain in registers. Only
// the weight gradients (errors x activations transposed, summed over samples)
// go through LDS, to transpose 16x16 tiles.
//
// Parameters (5,123), fp32 master copies, packed as
//   W1 [64][11], b1 [64], W2 [64][64], b2 [64], W3 [3][64], b3 [3].
// The input vector is (z0..z7, dx, dy, dz); b1 rides in W1's padded column
// 11 against a constant input 1.
#include <hip/hip_fp16.h>

#include <algorithm>

#include "gs3d_internal.h"

namespace gs3d {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// f32 -> bf16, round to nearest even: the hardware conversion
// (v_cvt_pk_bf16_f32, two values per instruction), the same bits as rounding
// the f32 pattern by hand for finite inputs.
__device__ __forceinline__ uint32_t bf16x2_bits(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t){a, b}, bf16x2_t));
}
__device__ __forceinline__ short bf16(float f) { return (short)(bf16x2_bits(f, 0.0f) & 0xFFFFu); }
__device__ __forceinline__ float bf16_to_f32(short h) { return __uint_as_float(((uint32_t)(uint16_t)h) << 16); }
__device__ __forceinline__ float f16_bits_to_f32(uint16_t h) { return __half2float(__ushort_as_half(h)); }
__device__ __forceinline__ uint16_t f32_to_f16_bits(float f) { return __half_as_ushort(__float2half_rn(f)); }

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    const uint32_t lo = bf16x2_bits(a, b), hi = bf16x2_bits(c, d);
    return __builtin_bit_cast(bf16x4, make_uint2(lo, hi));
}
// ReLU on the rounded bf16 patterns (v_pk_max_i16, two values per
// instruction): read as signed 16-bit integers, every negative value and -0
// is below 0 and the non-negative ones keep their order, and rounding
// commutes with max(x, 0), so the bits equal bf16(max(x, 0)).
typedef short s16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t bits) {
    const s16x2_t zero = {0, 0};
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2_t, bits), zero));
}
__device__ __forceinline__ bf16x4 pack_relu(f32x4 v) {
    const uint32_t lo = relu_bf16x2(bf16x2_bits(v[0], v[1])), hi = relu_bf16x2(bf16x2_bits(v[2], v[3]));
    return __builtin_bit_cast(bf16x4, make_uint2(lo, hi));
}

// Offsets into the packed parameter vector.
constexpr int kW1 = 0, kB1 = kW1 + 64 * 11, kW2 = kB1 + 64, kB2 = kW2 + 64 * 64, kW3 = kB2 + 64,
              kB3 = kW3 + 3 * 64;
static_assert(kB3 + 3 == GS3D_LEARNED_PARAMS, "parameter count");

// Padded W1 (64 x 16): columns 0..10 = W1, 11 = b1, 12..15 = 0.
__device__ __forceinline__ float w1p(const float* p, int i, int k) {
    return k < 11 ? p[kW1 + i * 11 + k] : (k == 11 ? p[kB1 + i] : 0.f);
}
// Padded W3 (16 x 64): rows 0..2 = W3.
__device__ __forceinline__ float w3p(const float* p, int c, int h) { return c < 3 ? p[kW3 + c * 64 + h] : 0.f; }

// The A-operand tiles of the forward pass, held in registers by every lane.
struct Forward {
    bf16x4 w1[4];      // W1p rows 16m.., K = 16
    bf16x4 w2[4][4];   // W2 [m][kk]
    bf16x4 w3[4];      // W3p, K tiles kk
    f32x4 b2[4], b3;   // accumulator initial values (D layout)
};

struct Backward {
    bf16x4 w3t[4];     // W3p^T (64 x 16): rows 16m.. hidden, K = output channel
    bf16x4 w2t[4][4];  // W2^T [m][kk]
    bf16x4 w1t[4];     // W1p^T (16 x 64): K tiles kk
};

__device__ __forceinline__ void load_forward(const float* p, Forward& f) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int i = 16 * m + r;
        f.w1[m] = pack4(w1p(p, i, 4 * g), w1p(p, i, 4 * g + 1), w1p(p, i, 4 * g + 2), w1p(p, i, 4 * g + 3));
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
            const float* row = p + kW2 + i * 64 + 16 * kk + 4 * g;
            f.w2[m][kk] = pack4(row[0], row[1], row[2], row[3]);
        }
        const int k0 = 16 * m + 4 * g;  // m plays kk for W3p
        f.w3[m] = pack4(w3p(p, r, k0), w3p(p, r, k0 + 1), w3p(p, r, k0 + 2), w3p(p, r, k0 + 3));
        const int o = 16 * m + 4 * g;
        f.b2[m] = f32x4{p[kB2 + o], p[kB2 + o + 1], p[kB2 + o + 2], p[kB2 + o + 3]};
    }
    f.b3 = f32x4{0.f, 0.f, 0.f, 0.f};
    if (g == 0) f.b3 = f32x4{p[kB3], p[kB3 + 1], p[kB3 + 2], 0.f};
}

__device__ __forceinline__ void load_backward(const float* p, Backward& b) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int h = 16 * m + r;  // hidden unit (row of the transposed matrix)
        b.w3t[m] = pack4(w3p(p, 4 * g, h), w3p(p, 4 * g + 1, h), w3p(p, 4 * g + 2, h), w3p(p, 4 * g + 3, h));
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
            const int o = 16 * kk + 4 * g;
            b.w2t[m][kk] = pack4(p[kW2 + o * 64 + h], p[kW2 + (o + 1) * 64 + h], p[kW2 + (o + 2) * 64 + h],
                                 p[kW2 + (o + 3) * 64 + h]);
        }
        const int o = 16 * m + 4 * g;  // m plays kk for W1p^T; row = input feature r
        b.w1t[m] = pack4(w1p(p, o, r), w1p(p, o + 1, r), w1p(p, o + 2, r), w1p(p, o + 3, r));
    }
}

// Activations of one 16-sample tile.
struct Acts {
    f32x4 d1[4], d2[4], d3;   // pre-activations (D layout)
    bf16x4 h1[4], h2[4];      // post-ReLU, bf16 (B layout)
};

__device__ __forceinline__ void forward(const Forward& f, bf16x4 x, Acts& a) {
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int m = 0; m < 4; m++) a.d1[m] = mfma(f.w1[m], x, zero);
#pragma unroll
    for (int m = 0; m < 4; m++) a.h1[m] = pack_relu(a.d1[m]);
#pragma unroll
    for (int m = 0; m < 4; m++) {
        f32x4 acc = f.b2[m];
#pragma unroll
        for (int kk = 0; kk < 4; kk++) acc = mfma(f.w2[m][kk], a.h1[kk], acc);
        a.d2[m] = acc;
    }
#pragma unroll
    for (int m = 0; m < 4; m++) a.h2[m] = pack_relu(a.d2[m]);
    f32x4 acc = f.b3;
#pragma unroll
    for (int kk = 0; kk < 4; kk++) acc = mfma(f.w3[kk], a.h2[kk], acc);
    a.d3 = acc;
}

// v_rcp_f32 (1 ulp) instead of an IEEE division: the results go to f16 colours
// or into an MSE loss, far coarser than the reciprocal's error
__device__ __forceinline__ float sigmoid(float x) { return __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }

// Input tile of 16 samples: lane (j = l % 16, g = l / 16) supplies features
// 4g..4g+3 of sample j: g 0, 1 the latent (as bf16), g 2 the direction and
// the constant 1 for b1, g 3 padding.
__device__ __forceinline__ bf16x4 input_tile(const float* z8, float dx, float dy, float dz) {
    const int g = (threadIdx.x & 63) >> 4;
    if (g == 0) return pack4(z8[0], z8[1], z8[2], z8[3]);
    if (g == 1) return pack4(z8[4], z8[5], z8[6], z8[7]);
    if (g == 2) return pack4(dx, dy, dz, 1.0f);
    return pack4(0.f, 0.f, 0.f, 0.f);
}

// ---- inference ------------------------------------------------------------------
// colors[v * n + i] = (r, g, b, 1) as f16 for every splat i and view v; the
// direction is normalize(pos_i - eye_v) as in the SH path. One wavefront takes
// 64 splats (4 tiles of 16) and loops over the views.
__global__ __launch_bounds__(256) void k_learned_colors(LearnedColorArgs a) {
    const int lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    const uint32_t wave = (blockIdx.x * 256u + threadIdx.x) >> 6;
    const uint32_t base = wave * 64u;
    if (base >= a.n) return;  // whole wavefronts exit together
    Forward f;
    load_forward(a.params, f);
    // Per tile, this lane's 4 input features that do not depend on the view
    // (input_tile's layout: g 0, 1 the latent as bf16, g 3 padding), packed
    // once; the g 2 lanes supply the direction per view. Only the position is
    // kept besides, so the view loop holds few registers.
    bf16x4 xin[4];
    float px[4], py[4], pz[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const uint32_t i = min(base + 16u * t + j, a.n - 1u);
        const uint2 h = reinterpret_cast<const uint2*>(a.latents)[2 * i + (g & 1)];  // halves 4(g&1)..+3
        float z4[4] = {f16_bits_to_f32((uint16_t)(h.x & 0xFFFFu)), f16_bits_to_f32((uint16_t)(h.x >> 16)),
                       f16_bits_to_f32((uint16_t)(h.y & 0xFFFFu)), f16_bits_to_f32((uint16_t)(h.y >> 16))};
        xin[t] = g < 2 ? pack4(z4[0], z4[1], z4[2], z4[3]) : pack4(0.f, 0.f, 0.f, 0.f);
        const float4 q = a.pos[i];
        px[t] = q.x, py[t] = q.y, pz[t] = q.z;
    }
    for (uint32_t v = 0; v < a.n_views; v++) {
        const float ex = a.eyes[4 * v], ey = a.eyes[4 * v + 1], ez = a.eyes[4 * v + 2];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const float dx = px[t] - ex, dy = py[t] - ey, dz = pz[t] - ez;
            const float inv = __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz);  // v_rsq_f32
            const bf16x4 x = g == 2 ? pack4(dx * inv, dy * inv, dz * inv, 1.0f) : xin[t];
            Acts act;
            forward(f, x, act);
            const uint32_t i = base + 16u * t + j;
            if (g == 0 && i < a.n) {
                const uint32_t rg = f32_to_f16_bits(sigmoid(act.d3[0])) | ((uint32_t)f32_to_f16_bits(sigmoid(act.d3[1])) << 16);
                const uint32_t ba = f32_to_f16_bits(sigmoid(act.d3[2])) | (0x3C00u << 16);
                a.colors[(size_t)v * a.n + i] = make_uint2(rg, ba);
            }
        }
    }
}

// ---- training -----------------------------------------------------------------------

// 16x16 tile transposes through LDS: a D-layout tile (lane (j, g) holds rows
// 4g..4g+3 of column j) is stored column-major, then read back as an A
// operand (rows = the D rows, K = the D columns) or a B operand (K = the D
// columns, N = the D rows).
struct TileLds {
    short v[16][16];  // [column][row]
};

__device__ __forceinline__ void store_tile(TileLds& t, f32x4 d) {
    const int lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    bf16x4 p = pack4(d[0], d[1], d[2], d[3]);
    *reinterpret_cast<bf16x4*>(&t.v[j][4 * g]) = p;
}
__device__ __forceinline__ void store_tile_bf16(TileLds& t, bf16x4 p) {
    const int lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    *reinterpret_cast<bf16x4*>(&t.v[j][4 * g]) = p;
}
// A[row r][k] = D[r][k]: lane (r = l % 16, g) reads D[r][4g + e].
__device__ __forceinline__ bf16x4 load_as_a(const TileLds& t) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    bf16x4 o;
#pragma unroll
    for (int e = 0; e < 4; e++) o[e] = t.v[4 * g + e][r];
    return o;
}
// B[k][n] = D[n][k]: lane (n = l % 16, g) reads D[n][4g + e].
__device__ __forceinline__ bf16x4 load_as_bt(const TileLds& t) { return load_as_a(t); }

// Per wavefront LDS: the tiles of one 16-sample step that the weight
// gradients need transposed.
struct WaveLds {
    TileLds x, e3;        // input, output error (16 rows each)
    TileLds h1[4], h2[4]; // activations (64 rows)
    TileLds e2[4], e1[4]; // errors of the hidden layers (64 rows)
};

constexpr int kTrainWaves = 4;

// One training step over splats [0, n) x directions dirs[first .. first + 16):
// forward, loss, backward; weight gradients summed per wavefront into
// grad_partial[wave][GS3D_LEARNED_PARAMS], latent gradients into
// grad_latents[i][8], loss partial sums into loss_partial[wave].
__global__ __launch_bounds__(kTrainWaves * 64) void k_learned_train(LearnedTrainArgs a) {
    __shared__ WaveLds s_lds[kTrainWaves];
    const int lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    const int wib = threadIdx.x >> 6;
    WaveLds& L = s_lds[wib];
    const uint32_t wave = blockIdx.x * kTrainWaves + wib;
    const uint32_t n_waves = gridDim.x * kTrainWaves;
    Forward f;
    Backward b;
    load_forward(a.params, f);
    load_backward(a.params, b);
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 gw1[4], gw2[4][4], gw3[4], gb2[4], gb3 = zero;
#pragma unroll
    for (int m = 0; m < 4; m++) {
        gw1[m] = zero, gw3[m] = zero, gb2[m] = zero;
#pragma unroll
        for (int kk = 0; kk < 4; kk++) gw2[m][kk] = zero;
    }
    float loss = 0.f;
    // this lane's direction and its SH basis (the same 16 for every splat of the step)
    const uint32_t dj = a.first_dir + j;
    const float dx = a.dirs[3 * dj], dy = a.dirs[3 * dj + 1], dz = a.dirs[3 * dj + 2];
    float basis[16];
#pragma unroll
    for (int k = 0; k < 16; k++) basis[k] = a.basis[16 * dj + k];
    const int ncoef = (a.sh_degree + 1) * (a.sh_degree + 1);

    // The next splat's SH coefficients and latent are loaded while the current
    // one is processed (one wavefront per SIMD leaves no other wave to hide
    // the load latency).
    float coef[16][3], zc[8];
    auto load_splat = [&](uint32_t i, float (*c)[3], float* z) {
#pragma unroll
        for (int k = 0; k < 16; k++)
#pragma unroll
            for (int ch = 0; ch < 3; ch++) c[k][ch] = k < ncoef ? a.sh[(size_t)(k * 3 + ch) * a.n + i] : 0.f;
        const float4* zp = reinterpret_cast<const float4*>(a.latents + (size_t)i * 8);
        const float4 z0 = zp[0], z1 = zp[1];
        z[0] = z0.x, z[1] = z0.y, z[2] = z0.z, z[3] = z0.w, z[4] = z1.x, z[5] = z1.y, z[6] = z1.z, z[7] = z1.w;
    };
    if (wave < a.n) load_splat(__builtin_amdgcn_readfirstlane(wave), coef, zc);
    for (uint32_t iv = wave; iv < a.n; iv += n_waves) {
        const uint32_t i = __builtin_amdgcn_readfirstlane(iv);  // one splat per wavefront
        float tgt[3], z8[8];
        // target: exact SH colour clamp01(0.5 + sum_k b_k c_k) for direction j
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < 16; k++) acc += basis[k] * coef[k][ch];
            const float v = acc + 0.5f;
            tgt[ch] = v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) z8[k] = f16_bits_to_f32(f32_to_f16_bits(zc[k]));
        if (iv + n_waves < a.n) load_splat(__builtin_amdgcn_readfirstlane(iv + n_waves), coef, zc);
        const bf16x4 x = input_tile(z8, dx, dy, dz);
        Acts act;
        forward(f, x, act);
        // output error (lanes g == 0 hold channels 0..2 of sample j)
        f32x4 e3 = zero;
        if (g == 0) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float y = sigmoid(act.d3[c]);
                const float diff = y - tgt[c];
                loss += diff * diff;
                e3[c] = a.grad_scale * 2.0f * diff * y * (1.0f - y);
            }
        }
        const bf16x4 e3b = pack4(e3[0], e3[1], e3[2], e3[3]);
        // hidden errors
        f32x4 e2[4], e1[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            f32x4 v = mfma(b.w3t[m], e3b, zero);
#pragma unroll
            for (int e = 0; e < 4; e++) v[e] = act.d2[m][e] > 0.f ? v[e] : 0.f;
            e2[m] = v;
        }
        bf16x4 e2b[4];
#pragma unroll
        for (int m = 0; m < 4; m++) e2b[m] = pack4(e2[m][0], e2[m][1], e2[m][2], e2[m][3]);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            f32x4 v = zero;
#pragma unroll
            for (int kk = 0; kk < 4; kk++) v = mfma(b.w2t[m][kk], e2b[kk], v);
#pragma unroll
            for (int e = 0; e < 4; e++) v[e] = act.d1[m][e] > 0.f ? v[e] : 0.f;
            e1[m] = v;
        }
        bf16x4 e1b[4];
#pragma unroll
        for (int m = 0; m < 4; m++) e1b[m] = pack4(e1[m][0], e1[m][1], e1[m][2], e1[m][3]);
        // input error -> latent gradient: rows 0..7, summed over the 16 directions
        f32x4 ex = zero;
#pragma unroll
        for (int kk = 0; kk < 4; kk++) ex = mfma(b.w1t[kk], e1b[kk], ex);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            float v = ex[e];
            v += __shfl_xor(v, 1, 64);
            v += __shfl_xor(v, 2, 64);
            v += __shfl_xor(v, 4, 64);
            v += __shfl_xor(v, 8, 64);
            ex[e] = v;
        }
        if (j == 0 && g < 2) {
#pragma unroll
            for (int e = 0; e < 4; e++) a.grad_latents[(size_t)i * 8 + 4 * g + e] = ex[e];
        }
        // bias gradients accumulate per lane (summed over lanes at the end)
#pragma unroll
        for (int m = 0; m < 4; m++) gb2[m] += e2[m];
        gb3 += e3;
        // weight gradients: errors (A, K = samples) x activations^T (B)
        store_tile_bf16(L.x, x);
        store_tile(L.e3, e3);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            store_tile_bf16(L.h1[m], act.h1[m]);
            store_tile_bf16(L.h2[m], act.h2[m]);
            store_tile_bf16(L.e2[m], e2b[m]);
            store_tile_bf16(L.e1[m], e1b[m]);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const bf16x4 xt = load_as_bt(L.x);
        const bf16x4 e3a = load_as_a(L.e3);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const bf16x4 e1a = load_as_a(L.e1[m]);
            gw1[m] = mfma(e1a, xt, gw1[m]);
            const bf16x4 e2a = load_as_a(L.e2[m]);
#pragma unroll
            for (int kk = 0; kk < 4; kk++) gw2[m][kk] = mfma(e2a, load_as_bt(L.h1[kk]), gw2[m][kk]);
            gw3[m] = mfma(e3a, load_as_bt(L.h2[m]), gw3[m]);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    // bias sums over the 16 sample lanes
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            float v = gb2[m][e];
            v += __shfl_xor(v, 1, 64);
            v += __shfl_xor(v, 2, 64);
            v += __shfl_xor(v, 4, 64);
            v += __shfl_xor(v, 8, 64);
            gb2[m][e] = v;
        }
#pragma unroll
    for (int e = 0; e < 4; e++) {
        float v = gb3[e];
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 8, 64);
        gb3[e] = v;
    }
    float* out = a.grad_partial + (size_t)wave * GS3D_LEARNED_PARAMS;
    // D layout of gradient tiles: lane (col = l % 16, g) holds rows 4g..4g+3
#pragma unroll
    for (int m = 0; m < 4; m++) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int row = 16 * m + 4 * g + e;  // hidden unit of layer 1
            if (j < 11) out[kW1 + row * 11 + j] = gw1[m][e];
            if (j == 11) out[kB1 + row] = gw1[m][e];
#pragma unroll
            for (int kk = 0; kk < 4; kk++) out[kW2 + row * 64 + 16 * kk + j] = gw2[m][kk][e];
            if (j == 0) out[kB2 + row] = gb2[m][e];
            const int c = 4 * g + e;  // output channel of layer 3
            if (c < 3) out[kW3 + c * 64 + 16 * m + j] = gw3[m][e];
        }
    }
    if (j == 0 && g == 0) {
        out[kB3] = gb3[0];
        out[kB3 + 1] = gb3[1];
        out[kB3 + 2] = gb3[2];
    }
    float l = loss;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if (lane == 0) a.loss_partial[wave] = l;
}

// Two-level sum over the per-wavefront rows: k_learned_reduce_rows folds row
// group y (rows y, y + kReduceGroups, ...) into fold[y][p]; k_learned_reduce
// sums the kReduceGroups folds. Both are coalesced over p.
constexpr uint32_t kReduceGroups = 64;

__global__ __launch_bounds__(256) void k_learned_reduce_rows(const float* partial, uint32_t rows, float* fold,
                                                            const float* loss_partial, float* loss_fold) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x, y = blockIdx.y;
    if (p < GS3D_LEARNED_PARAMS) {
        float s = 0.f;
        for (uint32_t r = y; r < rows; r += kReduceGroups) s += partial[(size_t)r * GS3D_LEARNED_PARAMS + p];
        fold[(size_t)y * GS3D_LEARNED_PARAMS + p] = s;
    }
    if (p == 0) {
        float s = 0.f;
        for (uint32_t r = y; r < rows; r += kReduceGroups) s += loss_partial[r];
        loss_fold[y] = s;
    }
}

__global__ __launch_bounds__(256) void k_learned_reduce(const float* fold, float* grad, const float* loss_fold,
                                                       float* loss) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p < GS3D_LEARNED_PARAMS) {
        float s = 0.f;
        for (uint32_t y = 0; y < kReduceGroups; y++) s += fold[(size_t)y * GS3D_LEARNED_PARAMS + p];
        grad[p] = s;
    }
    if (p == 0 && loss) {
        float s = 0.f;
        for (uint32_t y = 0; y < kReduceGroups; y++) s += loss_fold[y];
        *loss = s;
    }
}

// Adam (Kingma & Ba 2015) on fp32 masters: m = b1 m + (1-b1) g, v = b2 v +
// (1-b2) g^2, x -= lr * mhat / (sqrt(vhat) + eps), bias-corrected by step t.
__global__ __launch_bounds__(256) void k_adam(float* x, float* m, float* v, const float* grad, size_t n,
                                              float lr, float b1, float b2, float eps, float c1, float c2,
                                              uint16_t* f16_copy) {
    for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (size_t)gridDim.x * 256u) {
        const float gi = grad[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float xi = x[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        x[i] = xi;
        if (f16_copy) f16_copy[i] = f32_to_f16_bits(xi);
    }
}

__global__ __launch_bounds__(256) void k_f16_copy(const float* x, size_t n, uint16_t* f16_copy) {
    for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (size_t)gridDim.x * 256u)
        f16_copy[i] = f32_to_f16_bits(x[i]);
}

// *out += sum over pixels of (a - b)^2 over the RGB channels of f16 RGBA images.
__global__ __launch_bounds__(256) void k_sq_diff_sum(const uint16_t* a, const uint16_t* b, size_t pixels,
                                                    double* out) {
    __shared__ float s_part[256 / 64];
    float acc = 0.f;
    for (size_t p = (size_t)blockIdx.x * 256u + threadIdx.x; p < pixels; p += (size_t)gridDim.x * 256u) {
        const uint2 x = reinterpret_cast<const uint2*>(a)[p];
        const uint2 y = reinterpret_cast<const uint2*>(b)[p];
        const float d0 = f16_bits_to_f32((uint16_t)(x.x & 0xFFFFu)) - f16_bits_to_f32((uint16_t)(y.x & 0xFFFFu));
        const float d1 = f16_bits_to_f32((uint16_t)(x.x >> 16)) - f16_bits_to_f32((uint16_t)(y.x >> 16));
        const float d2 = f16_bits_to_f32((uint16_t)(x.y & 0xFFFFu)) - f16_bits_to_f32((uint16_t)(y.y & 0xFFFFu));
        acc += d0 * d0 + d1 * d1 + d2 * d2;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int w = 0; w < 256 / 64; w++) t += s_part[w];
        atomicAdd(out, (double)t);
    }
}

}  // namespace

hipError_t launch_sq_diff_sum(const uint16_t* a, const uint16_t* b, size_t pixels, double* out, hipStream_t s) {
    if (pixels == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((pixels + 255) / 256, 2048);
    hipLaunchKernelGGL(k_sq_diff_sum, dim3((uint32_t)blocks), dim3(256), 0, s, a, b, pixels, out);
    return hipGetLastError();
}

uint32_t learned_train_waves(uint32_t n) {
    const uint32_t want = (n + 63u) / 64u;  // at least 64 splats per wavefront
    const uint32_t w = std::max<uint32_t>(kTrainWaves, std::min(4096u, want));
    return (w + kTrainWaves - 1) / kTrainWaves * kTrainWaves;
}

hipError_t launch_learned_colors(const LearnedColorArgs& a, hipStream_t s) {
    if (a.n == 0 || a.n_views == 0) return hipSuccess;
    const uint32_t waves = (a.n + 63u) / 64u;
    hipLaunchKernelGGL(k_learned_colors, dim3((waves + 3u) / 4u), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_learned_train(const LearnedTrainArgs& a, uint32_t waves, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    if (waves % kTrainWaves) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_learned_train, dim3(waves / kTrainWaves), dim3(kTrainWaves * 64), 0, s, a);
    return hipGetLastError();
}

size_t learned_reduce_scratch_floats() { return (size_t)kReduceGroups * (GS3D_LEARNED_PARAMS + 1); }

hipError_t launch_learned_reduce(const float* partial, uint32_t rows, float* grad, const float* loss_partial,
                                 float* loss, float* scratch, hipStream_t s) {
    float* fold = scratch;
    float* loss_fold = scratch + (size_t)kReduceGroups * GS3D_LEARNED_PARAMS;
    const uint32_t blocks = (GS3D_LEARNED_PARAMS + 255) / 256;
    hipLaunchKernelGGL(k_learned_reduce_rows, dim3(blocks, kReduceGroups), dim3(256), 0, s, partial, rows, fold,
                       loss_partial, loss_fold);
    hipLaunchKernelGGL(k_learned_reduce, dim3(blocks), dim3(256), 0, s, fold, grad, loss_fold, loss);
    return hipGetLastError();
}

hipError_t launch_adam(float* x, float* m, float* v, const float* grad, size_t n, float lr, float b1, float b2,
                       float eps, uint32_t step, uint16_t* f16_copy, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_adam, dim3((uint32_t)blocks), dim3(256), 0, s, x, m, v, grad, n, lr, b1, b2, eps, c1, c2,
                       f16_copy);
    return hipGetLastError();
}

hipError_t launch_f16_copy(const float* x, size_t n, uint16_t* f16_copy, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_f16_copy, dim3((uint32_t)blocks), dim3(256), 0, s, x, n, f16_copy);
    return hipGetLastError();
}

}  // namespace gs3d
