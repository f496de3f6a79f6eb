/*
 * gs3d_oracle.c — CPU oracle of the rebin-then-rasterize path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it (through
 * ctypes, as oracle/liboracle_gs3d.so). The product library never links it.
 *
 * It restates, in plain C, the reference WGSL pipeline (paths relative to the
 * reference repo):
 *   preprocess   src/shaders/compute/preprocess_splats.wgsl:49-184
 *   SH colour    (new; no reference function) real SH basis of degree 0-3 as
 *                published with 3DGS (Kerbl et al. 2023), reducing to the
 *                reference DC colour clamp01(0.5 + C0 f_dc), SceneWorker.ts:428-433
 *   scan         prefix_scan_local/blocks/add.wgsl + scan_core.wgsl: an
 *                exclusive u32 sum mod 2^32 with total sentinel at [N]
 *   emit         emit_tile_refs.wgsl:56-79 (row-major tile loop)
 *   sort         radix_histogram.wgsl + scan_local/scan_add.wgsl +
 *                radix_scatter.wgsl: stable LSD counting sort, 8-bit digits
 *                over 64 bits (GaussianSplatRenderer.ts:88-91, 1216-1247)
 *   identify     identify_tile_ranges.wgsl:30-52 (trailing tiles stay 0)
 *   rasterize    rasterize_tiles.wgsl:43-110 (accumulate, then stop if T<1e-4)
 *   fp16 store   rgba16float target (GaussianSplatRenderer.ts:233-238), RNE
 *   heat map     debug_tile_overlay.wgsl:35-58 (tint over the target, f16 store)
 *   composite    composite_splats.wgsl:25-29 with the premultiplied "over"
 *                blend state of GaussianSplatRenderer.ts:939-945, unorm8 store
 *
 * The body in gs3d_oracle_real.inc is compiled twice: REAL=float (the fp32
 * mirror: operation order of the WGSL, no FMA contraction — built with
 * -ffp-contract=off) and REAL=double (the host FP64 render).
 *
 * Pinning (the tests that check each function): scan vs numpy cumsum incl.
 * mod-2^32 wrap; emit vs a python enumeration; sort vs np.lexsort and brute
 * force; identify vs the closed form; preprocess vs closed-form KATs
 * (centred isotropic splat, culling cases, radius floor, tile clamp), scipy
 * Rotation for the quaternion matrix, eigenvalues of cov3d; SH basis vs
 * scipy.special.sph_harm_y; raster vs single-splat and occlusion KATs and a
 * brute-force per-pixel composite; fp16 vs numpy float16; heat map vs its
 * colour-ramp end points (green, yellow, red), the underflow guard and the
 * clamp of debug_ref; composite vs the "over" identities (transparent source
 * keeps the framebuffer, opaque source replaces it, associativity) and a
 * float64 blend rounded to unorm8.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/gs3d.h"

static int g_threads = 1;

/* Number of OpenMP threads the oracle uses (1 = scalar). Returns the value set. */
int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n < 1) n = 1;
    g_threads = n;
    omp_set_num_threads(n);
#else
    (void)n;
    g_threads = 1;
#endif
    return g_threads;
}

/* ---- fp16 (IEEE binary16, round to nearest even) ------------------------ */
uint16_t orc_f32_to_f16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xffu) {                                  /* inf / nan */
        return (uint16_t)(sign | 0x7c00u | (man ? 0x200u : 0u));
    }
    int e = (int)exp - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);      /* overflow -> inf */
    if (e <= 0) {                                        /* subnormal / zero */
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        int shift = 14 - e;                              /* 1 - e + 13 */
        uint32_t half_man = man >> shift;
        uint32_t rem = man & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (half_man & 1u))) half_man++;
        return (uint16_t)(sign | half_man);
    }
    uint32_t h = sign | ((uint32_t)e << 10) | (man >> 13);
    uint32_t rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;  /* may carry into exp: correct */
    return (uint16_t)h;
}

float orc_f16_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t exp = ((uint32_t)h >> 10) & 0x1fu;
    uint32_t man = (uint32_t)h & 0x3ffu;
    uint32_t x;
    if (exp == 0) {
        if (man == 0) {
            x = sign;
        } else {                                         /* normalise subnormal */
            int e = -1;
            do { e++; man <<= 1; } while (!(man & 0x400u));
            x = sign | ((uint32_t)(127 - 15 - e) << 23) | ((man & 0x3ffu) << 13);
        }
    } else if (exp == 31) {
        x = sign | 0x7f800000u | (man << 13);
    } else {
        x = sign | ((exp + 127 - 15) << 23) | (man << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

void orc_f32_to_f16_array(const float* in, uint16_t* out, size_t n) {
    for (size_t i = 0; i < n; i++) out[i] = orc_f32_to_f16(in[i]);
}

/* ---- scan (prefix_scan_*.wgsl) ----------------------------------------- */
/* offsets[i] = sum_{j<i} counts[j] (u32 wrap), offsets[n] = ref_counter = total. */
uint32_t orc_scan(const uint32_t* counts, uint32_t n, uint32_t* offsets) {
    uint32_t running = 0;
    for (uint32_t i = 0; i < n; i++) {
        offsets[i] = running;
        running += counts[i];
    }
    offsets[n] = running;
    return running;
}

/* ---- emit (emit_tile_refs.wgsl:56-79) ----------------------------------- */
void orc_emit(const gs3d_projected_splat* ps, const uint32_t* offsets, uint32_t n,
              uint32_t tiles_x, uint64_t* keys, uint32_t* values) {
    for (uint32_t id = 0; id < n; id++) {
        const gs3d_projected_splat* s = &ps[id];
        if (s->valid == 0u) continue;
        float neg = -s->depth;
        uint32_t lo;
        memcpy(&lo, &neg, 4);
        uint32_t w = offsets[id];
        for (uint32_t ty = s->tile_min[1]; ty <= s->tile_max[1]; ty++) {
            for (uint32_t tx = s->tile_min[0]; tx <= s->tile_max[0]; tx++) {
                uint32_t tile = ty * tiles_x + tx;
                keys[w] = ((uint64_t)tile << 32) | (uint64_t)lo;
                values[w] = id;
                w++;
            }
        }
    }
}

/* ---- sort (radix_*.wgsl): stable LSD counting sort ---------------------- */
/* One pass = stable partition by the digit at `shift`. With g_threads > 1 the
 * input is split into contiguous chunks whose per-digit offsets are laid out
 * chunk-major inside each digit, which keeps the pass stable. */
static void orc_sort_pass(const uint64_t* kin, const uint32_t* vin, uint64_t* kout,
                          uint32_t* vout, size_t n, int shift, int bits) {
    const uint32_t buckets = 1u << bits;
    const uint64_t mask = buckets - 1u;
    int T = g_threads;
    if ((size_t)T > n / 4096 + 1) T = (int)(n / 4096 + 1);
    size_t* hist = (size_t*)calloc((size_t)T * buckets, sizeof(size_t));
    size_t chunk = (n + (size_t)T - 1) / (size_t)T;
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < T; t++) {
        size_t lo = (size_t)t * chunk, hi = lo + chunk < n ? lo + chunk : n;
        size_t* h = hist + (size_t)t * buckets;
        for (size_t i = lo; i < hi; i++) h[(kin[i] >> shift) & mask]++;
    }
    size_t run = 0;
    for (uint32_t d = 0; d < buckets; d++) {
        for (int t = 0; t < T; t++) {
            size_t c = hist[(size_t)t * buckets + d];
            hist[(size_t)t * buckets + d] = run;
            run += c;
        }
    }
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (int t = 0; t < T; t++) {
        size_t lo = (size_t)t * chunk, hi = lo + chunk < n ? lo + chunk : n;
        size_t* h = hist + (size_t)t * buckets;
        for (size_t i = lo; i < hi; i++) {
            size_t p = h[(kin[i] >> shift) & mask]++;
            kout[p] = kin[i];
            vout[p] = vin[i];
        }
    }
    free(hist);
}

/* Sort n pairs in place over all 64 key bits with `bits`-wide digits (the
 * reference runs 8 passes of 8 bits, ending in the _a buffers). */
int orc_radix_sort(uint64_t* keys, uint32_t* values, size_t n, int bits) {
    if (bits != 4 && bits != 8) return -1;
    uint64_t* kb = (uint64_t*)malloc(n * sizeof(uint64_t) + 8);
    uint32_t* vb = (uint32_t*)malloc(n * sizeof(uint32_t) + 4);
    if (!kb || !vb) { free(kb); free(vb); return -1; }
    uint64_t* ka = keys; uint32_t* va = values;
    for (int shift = 0; shift < 64; shift += bits) {
        orc_sort_pass(ka, va, kb, vb, n, shift, bits);
        uint64_t* tk = ka; ka = kb; kb = tk;
        uint32_t* tv = va; va = vb; vb = tv;
    }
    /* 64/bits passes is even for 4 and 8: the result is back in keys/values. */
    if (ka != keys) { memcpy(keys, ka, n * 8); memcpy(values, va, n * 4); kb = ka; vb = va; }
    free(kb); free(vb);
    return 0;
}

/* ---- identify (identify_tile_ranges.wgsl:30-52) ------------------------- */
/* tile_offsets has tile_count+1 entries and is cleared here first
 * (GaussianSplatRenderer.ts:1137). */
void orc_identify(const uint64_t* keys, uint32_t n, uint32_t tile_count, uint32_t* tile_offsets) {
    memset(tile_offsets, 0, ((size_t)tile_count + 1) * 4);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t curr = (uint32_t)(keys[i] >> 32);
        if (i == 0 || (uint32_t)(keys[i - 1] >> 32) != curr) {
            uint32_t from = (i == 0) ? 0u : (uint32_t)(keys[i - 1] >> 32) + 1u;
            for (uint32_t t = from; t <= curr; t++) tile_offsets[t] = i;
        }
        if (i == n - 1 || (uint32_t)(keys[i + 1] >> 32) != curr) {
            tile_offsets[curr + 1] = i + 1;
        }
    }
}

/* ---- precision-templated stages ----------------------------------------- */
#define REAL float
#define RSUF(name) name##_f32
#define RSQRT sqrtf
#define RCEIL ceilf
#define RFLOOR floorf
#define REXP expf
#define RLIT(x) x##f
#include "gs3d_oracle_real.inc"
#undef REAL
#undef RSUF
#undef RSQRT
#undef RCEIL
#undef RFLOOR
#undef REXP
#undef RLIT

#define REAL double
#define RSUF(name) name##_f64
#define RSQRT sqrt
#define RCEIL ceil
#define RFLOOR floor
#define REXP exp
#define RLIT(x) x
#include "gs3d_oracle_real.inc"

/* ---- exported f32 stage wrappers (gs3d_projected_splat layout) ---------- */
/* orc_proj_t_f32 and gs3d_projected_splat have the same 16 x 4-byte layout. */
_Static_assert(sizeof(orc_proj_t_f32) == sizeof(gs3d_projected_splat), "projected layout");

void orc_preprocess_f32(const gs3d_camera* cam, uint32_t n, const float* pos4, const float* scale4,
                        const float* rot4, const float* color4, const float* f_dc,
                        const float* sh_rest, int deg, uint32_t tiles_x, uint32_t tiles_y,
                        float inv_tx, float inv_ty, gs3d_projected_splat* out, uint32_t* counts) {
    orc_preprocess_all_f32(cam, n, pos4, scale4, rot4, color4, f_dc, sh_rest, deg, tiles_x, tiles_y,
                           inv_tx, inv_ty, (orc_proj_t_f32*)out, counts);
}

/* rasterize_tiles.wgsl over caller buffers. With quantize_f16 the image is
 * rounded through binary16 (the rgba16float target). */
void orc_rasterize_f32(const gs3d_projected_splat* ps, const uint32_t* values,
                       const uint32_t* tile_offsets, uint32_t tile_x, uint32_t tile_y, uint32_t W,
                       uint32_t H, int quantize_f16, float* image, uint64_t* visited,
                       uint64_t* blended) {
    orc_raster_all_f32((const orc_proj_t_f32*)ps, values, tile_offsets, tile_x, tile_y, W, H, image,
                       visited, blended);
    if (quantize_f16) {
        size_t m = (size_t)W * H * 4;
        for (size_t i = 0; i < m; i++) image[i] = orc_f16_to_f32(orc_f32_to_f16(image[i]));
    }
}

/* ---- debug heat map (debug_tile_overlay.wgsl:35-58) ---------------------- */
/* base: W*H*4 premultiplied target (f16 values widened to f32); out: the
 * overlay, rounded through binary16 (the rgba16float debug target). The pixel
 * centre's tile is min(px / tile, tiles - 1) (:36-38); the count keeps the
 * underflow guard of :42-44; opacity 0.4 (:52). */
static float orc_clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

void orc_tile_overlay_f32(const float* base, const uint32_t* tile_offsets, uint32_t tile_x,
                          uint32_t tile_y, uint32_t W, uint32_t H, uint32_t debug_ref, float* out) {
    const uint32_t tiles_x = (W + tile_x - 1) / tile_x, tiles_y = (H + tile_y - 1) / tile_y;
    const float ref = (float)debug_ref > 1.0f ? (float)debug_ref : 1.0f;
    const float ao = 0.4f, keep = 1.0f - ao;
    for (uint32_t y = 0; y < H; y++) {
        for (uint32_t x = 0; x < W; x++) {
            uint32_t tx = x / tile_x, ty = y / tile_y;
            if (tx > tiles_x - 1) tx = tiles_x - 1;
            if (ty > tiles_y - 1) ty = tiles_y - 1;
            const uint32_t tile = ty * tiles_x + tx;
            const uint32_t lo = tile_offsets[tile], hi = tile_offsets[tile + 1];
            const uint32_t count = hi >= lo ? hi - lo : 0u;
            const float t = orc_clamp01((float)count / ref);
            const float tint[3] = {orc_clamp01(t * 2.0f), orc_clamp01(2.0f - t * 2.0f), 0.0f};
            const float* b = base + ((size_t)y * W + x) * 4;
            float* o = out + ((size_t)y * W + x) * 4;
            for (int ch = 0; ch < 3; ch++) {
                const float a = tint[ch] * ao, m = b[ch] * keep;
                o[ch] = orc_f16_to_f32(orc_f32_to_f16(a + m));
            }
            const float m = b[3] * keep;
            o[3] = orc_f16_to_f32(orc_f32_to_f16(ao + m));
        }
    }
}

/* ---- composite (composite_splats.wgsl:25-29, blend GaussianSplatRenderer.ts:939-945) */
/* fb: W*H*4 unorm8 bytes, R,G,B,A order (B,G,R,A when bgra), read unless
 * clear_rgba is set (then every pixel starts as the clear colour stored to
 * unorm8, the clear pass), and written. Per channel:
 * out = src + dst * (1 - src.a), then clamp to [0, 1], *255, round to nearest. */
static uint8_t orc_unorm8(float v) {
    v = v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f;
    const float s = v * 255.0f;
    return (uint8_t)(uint32_t)(s + 0.5f);
}

void orc_composite_f32(const float* src, uint32_t W, uint32_t H, int bgra, const float* clear_rgba,
                       uint8_t* fb) {
    const int ri = bgra ? 2 : 0, bi = bgra ? 0 : 2;
    uint8_t clear[4] = {0, 0, 0, 0};
    if (clear_rgba) {
        clear[ri] = orc_unorm8(clear_rgba[0]);
        clear[1] = orc_unorm8(clear_rgba[1]);
        clear[bi] = orc_unorm8(clear_rgba[2]);
        clear[3] = orc_unorm8(clear_rgba[3]);
    }
    const size_t px = (size_t)W * H;
    for (size_t p = 0; p < px; p++) {
        uint8_t* d = fb + p * 4;
        const uint8_t* in = clear_rgba ? clear : d;
        const float dst[4] = {(float)in[ri] / 255.0f, (float)in[1] / 255.0f, (float)in[bi] / 255.0f,
                              (float)in[3] / 255.0f};
        const float* s4 = src + p * 4;
        const float inv_a = 1.0f - s4[3];
        uint8_t o[4];
        for (int ch = 0; ch < 4; ch++) {
            const float m = dst[ch] * inv_a;
            o[ch] = orc_unorm8(s4[ch] + m);
        }
        d[ri] = o[0];
        d[1] = o[1];
        d[bi] = o[2];
        d[3] = o[3];
    }
}
