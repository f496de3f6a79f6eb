// k_raster.hip — tile ranges and the tiled front-to-back rasterizer.
//
// identify: the reference (identify_tile_ranges.wgsl:30-52) runs one thread
// per sorted key; at a tile START it fills tile_offsets[prev+1 .. curr] = i
// serially, at a tile END it writes tile_offsets[curr+1] = i+1, and tiles past
// the last populated one keep the 0 of the per-rebin clear. That array equals,
// for every tile t: lower_bound(key.hi >= t) if t <= last+1, else 0 (last =
// tile of the final key; all zero when there are no keys). One thread per tile
// evaluates that closed form with a search over the sorted keys (8-ary rounds,
// then binary), so no thread walks a long gap serially.
//
// rasterize: same per-pixel arithmetic and order as rasterize_tiles.wgsl:43-110
// (pixel centre +0.5; skip if |d|^2 > r^2; power = -1/2 d^T conic d, skip if
// power > 0; alpha = min(0.99, o e^power), skip if alpha < 1/255; C += T a c,
// T *= 1-a, THEN stop when T < 1e-4; out = (C, 1-T) as rgba16float). The
// schedule is CDNA-native: each wavefront owns a block of the tile's pixels
// (8 x 8 at one pixel per lane) and walks the tile's sorted list on its own,
// 64 splats per step: one lane gathers one splat (coalesced id load + three
// 16-B loads), splats that cannot reach the block are dropped by a ballot, and
// the survivors are compacted into the wavefront's LDS slots and broadcast from
// there. No workgroup barriers, and a wavefront leaves as soon as all its
// pixels are saturated. Splats that cannot reach alpha >= 1/255 inside the
// block (radius circle and the exact minimum of the conic over the block, see
// may_touch) are skipped before the broadcast; this drops no blend the
// per-pixel rules would keep.
// exp is the hardware exp2 (v_exp_f32) of the power pre-scaled by log2(e);
// results stay within one fp16 step of the fp32 restatement on the CPU.
#include <hip/hip_fp16.h>

#include <algorithm>

#include "gs3d_internal.h"

namespace gs3d {

namespace {

constexpr uint32_t kWideSearchKeys = 32u << 20;

// The tile id of sorted key i sits at word stride * i + word (64-bit keys:
// stride 2, the high 32-bit word; tile ids alone, W of 32 or 16 bits: stride 1).
template <typename W>
__global__ __launch_bounds__(256) void k_identify(const W* __restrict__ kw, uint32_t stride,
                                                  uint32_t word, const uint32_t* __restrict__ n_dev,
                                                  uint32_t* __restrict__ tile_offsets,
                                                  uint32_t tile_count, uint32_t n_max,
                                                  uint32_t* __restrict__ k_host) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t > tile_count) return;
    uint32_t n = *n_dev;
    if (t == 0 && k_host) *k_host = n;  // K for the host, no copy packet needed
    if (n > n_max) n = 0;  // keys did not fit: draw nothing, the host re-runs the frame
    uint32_t out = 0;
    if (n > 0) {
        const uint32_t last = (uint32_t)kw[(size_t)stride * (n - 1) + word];
        if (t <= last + 1u) {
            // first index whose tile id >= t, in [lo, hi]. While the range is
            // wide, 7 independent probes p_j = lo + j*step (all < hi <= n) cut it
            // about 8-fold per round: the c probes below t are a prefix, so the
            // answer lies in (p_c, p_{c+1}]. Then a binary search. Past
            // kWideSearchKeys keys the extra probes miss the cache and the
            // binary search alone is faster (measured: 3.6M keys 10.5 -> 9.3 us,
            // 118.9M keys 19.5 -> 23 us with the probes).
            uint32_t lo = 0, hi = n;
            while (n <= kWideSearchKeys && hi - lo > 8u) {
                const uint32_t step = (hi - lo) >> 3;
                uint32_t c = 0;
#pragma unroll
                for (uint32_t j = 1; j < 8; j++)
                    c += (uint32_t)kw[(size_t)stride * (lo + j * step) + word] < t ? 1u : 0u;
                const uint32_t nlo = c ? lo + c * step + 1u : lo;
                hi = c < 7u ? lo + (c + 1u) * step : hi;
                lo = nlo;
            }
            while (lo < hi) {
                const uint32_t mid = lo + ((hi - lo) >> 1);
                if ((uint32_t)kw[(size_t)stride * mid + word] < t) lo = mid + 1; else hi = mid;
            }
            out = lo;
        }
    }
    tile_offsets[t] = out;
}

template <typename W>
__global__ __launch_bounds__(256) void k_rebuild_keys(const W* __restrict__ tiles,
                                                      const uint32_t* __restrict__ values,
                                                      const gs3d_projected_splat* __restrict__ projected,
                                                      const uint32_t* __restrict__ n_dev, uint32_t n_max,
                                                      uint64_t* __restrict__ keys) {
    const uint32_t n = min(*n_dev, n_max);
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const float depth = projected[values[i]].depth;
        keys[i] = ((uint64_t)tiles[i] << 32) | (uint64_t)__float_as_uint(-depth);
    }
}

// Conservative test of one gathered splat against a wavefront's rectangle of
// pixel centres [rx0, rx1] x [ry0, ry1]. A pixel can only be blended when it is
// inside the radius circle AND alpha = o e^{-q/2} >= 1/255, i.e. when
// q = d^T conic d <= t = 2 ln(255 o). The circle test uses the distance from
// the mean to the rectangle. For the ellipse, q is convex, so when the mean is
// outside the rectangle the minimum of q over it lies on an edge facing the
// mean; on such an edge q is a 1-D parabola whose minimum is the clamped
// vertex. Both facing edges are evaluated (an edge that does not face gives a
// value at a point of the rectangle, never below the minimum), so qmin is the
// exact minimum over the rectangle. Both bounds are widened by far more than
// float rounding (1-2 % + 0.05), so the test only drops splats that cannot
// pass the exact per-pixel rules at any pixel centre of the rectangle.
__device__ __forceinline__ bool may_touch(float4 f0, float4 f1, float rx0, float rx1, float ry0,
                                          float ry1) {
    const float mx = f0.x, my = f0.y, rad = f0.w;
    const float ca = f1.x, cb = f1.y, cc = f1.z, op = f1.w;
    if (!(op * 255.0f >= 0.99f)) return false;  // alpha <= o < 1/255 everywhere
    const float qx = fminf(fmaxf(mx, rx0), rx1);
    const float qy = fminf(fmaxf(my, ry0), ry1);
    const float ex = qx - mx, ey = qy - my;
    const float rr = rad * 1.01f + 0.05f;
    if (ex * ex + ey * ey > rr * rr) return false;
    if (!(ca > 0.0f && cc > 0.0f && ca * cc - cb * cb > 0.0f)) return true;  // degenerate: keep
    // The vertex positions and the threshold use the hardware reciprocal and
    // log2 (about 1 ulp): a vertex off by that much moves q by far less than
    // the margins, and q at any point of the rectangle is >= its minimum.
    // edge x = qx: vertex at dy = -cb ex / cc, clamped to the rectangle
    const float vy = fminf(fmaxf(my - cb * ex * __builtin_amdgcn_rcpf(cc), ry0), ry1) - my;
    const float q1 = ca * ex * ex + 2.0f * cb * ex * vy + cc * vy * vy;
    // edge y = qy: vertex at dx = -cb ey / ca, clamped
    const float vx = fminf(fmaxf(mx - cb * ey * __builtin_amdgcn_rcpf(ca), rx0), rx1) - mx;
    const float q2 = ca * vx * vx + 2.0f * cb * vx * ey + cc * ey * ey;
    // 2 ln(255 o) = 2 ln(2) log2(255 o); 255 o >= 0.99 here, a normal number
    constexpr float kTwoLn2 = 1.3862943611198906f;
    const float t = kTwoLn2 * __builtin_amdgcn_logf(255.0f * op) * 1.02f + 0.05f;
    return fminf(q1, q2) * 0.99f <= t;
}

// may_touch for the four 4 x 4 quarters of an 8 x 8 block at once. Quarter
// q = 2 r + c covers the pixel centres x0 + 4 c + [0, 3], y0 + 4 r + [0, 3];
// hit[q] is may_touch on that rectangle, with the same tests and margins.
// Terms that depend on one column or one row only are formed once; a clamp
// is v_med3 (the bounds are ordered, the mean finite).
__device__ __forceinline__ void may_touch_quarters(float4 f0, float4 f1, float x0, float y0, bool hit[4]) {
    const float mx = f0.x, my = f0.y, rad = f0.w;
    const float ca = f1.x, cb = f1.y, cc = f1.z, op = f1.w;
    const bool visible = op * 255.0f >= 0.99f;
    const float rr = rad * 1.01f + 0.05f;
    const float rr2 = rr * rr;
    const bool conic_ok = ca > 0.0f && cc > 0.0f && ca * cc - cb * cb > 0.0f;
    const float kx = cb * __builtin_amdgcn_rcpf(cc);  // edge x = const: vertex at dy = -kx dx
    const float ky = cb * __builtin_amdgcn_rcpf(ca);  // edge y = const: vertex at dx = -ky dy
    const float cb2 = 2.0f * cb;
    // q_min * 0.99 <= t  as  q_min <= t / 0.99
    constexpr float kTwoLn2 = 1.3862943611198906f;
    const float t = (kTwoLn2 * __builtin_amdgcn_logf(255.0f * op) * 1.02f + 0.05f) * (1.0f / 0.99f);
    float lo_x[2], hi_x[2], lo_y[2], hi_y[2];
    float ex[2], ex2[2], ey[2], ey2[2];  // mean to the nearest column / row, squared
    float vy0[2], ax[2], bx[2];          // per column: edge-x vertex before its clamp, ca ex^2, 2 cb ex
    float vx0[2], cy[2], by[2];          // per row:    edge-y vertex before its clamp, cc ey^2, 2 cb ey
#pragma unroll
    for (int i = 0; i < 2; i++) {
        lo_x[i] = x0 + 4.0f * (float)i;
        hi_x[i] = lo_x[i] + 3.0f;
        lo_y[i] = y0 + 4.0f * (float)i;
        hi_y[i] = lo_y[i] + 3.0f;
        ex[i] = __builtin_amdgcn_fmed3f(mx, lo_x[i], hi_x[i]) - mx;
        ey[i] = __builtin_amdgcn_fmed3f(my, lo_y[i], hi_y[i]) - my;
        ex2[i] = ex[i] * ex[i];
        ey2[i] = ey[i] * ey[i];
        vy0[i] = fmaf(-kx, ex[i], my);
        vx0[i] = fmaf(-ky, ey[i], mx);
        ax[i] = ca * ex2[i];
        bx[i] = cb2 * ex[i];
        cy[i] = cc * ey2[i];
        by[i] = cb2 * ey[i];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int c = q & 1, r = q >> 1;
        const bool circle = ex2[c] + ey2[r] <= rr2;
        // edge x = mean's nearest column: q = ca ex^2 + 2 cb ex vy + cc vy^2
        const float vy = __builtin_amdgcn_fmed3f(vy0[c], lo_y[r], hi_y[r]) - my;
        const float q1 = fmaf(cc * vy, vy, fmaf(bx[c], vy, ax[c]));
        // edge y = mean's nearest row: q = ca vx^2 + 2 cb vx ey + cc ey^2
        const float vx = __builtin_amdgcn_fmed3f(vx0[r], lo_x[c], hi_x[c]) - mx;
        const float q2 = fmaf(ca * vx, vx, fmaf(by[r], vx, cy[r]));
        // a degenerate conic keeps every quarter its circle reaches
        hit[q] = visible & circle & (!conic_ok | (fminf(q1, q2) <= t));
    }
}

// Pixel block of one wavefront: BW columns x BH rows, each lane owning PPL
// pixels of one column spaced 64/BW rows apart. 8-pixel-wide blocks (8 x 8 at
// one pixel per lane) touch fewer splats than full-width rows of the same area.
template <int TW, int PPL>
struct RasterShape {
    static constexpr int BW = (8 * PPL <= TW) ? 8 : 16;
    static constexpr int BH = kWave * PPL / BW;
    static constexpr int kLaneStride = kWave / BW;  // rows between a lane's pixels
    static constexpr int kBlocksX = TW / BW;
    static constexpr int kWaves = (TW / BW) * (TW / BH);
    static_assert(BH <= TW && TW % BW == 0 && TW % BH == 0, "blocks tile the tile");
};

// One survivor as the composite loop reads it: {mean.x, mean.y, r^2, A},
// {B, C, opacity, red}, {green, blue} with A, B, C the conic scaled so that
// A dx^2 + B dx dy + C dy^2 = log2(e) * power, the argument of v_exp_f32.
struct alignas(16) Survivor {
    float4 a;
    float4 b;
    float2 c;
    float2 pad;
};

// One wavefront's LDS. Every form holds the batch's compacted survivors; the
// one-pixel-per-lane form also holds a zero-opacity record (slot kWave) and,
// per 16-lane quarter, the LDS addresses of the survivors that may touch the
// quarter's 4 x 4 pixels, padded with the zero record's address.
template <int PPL>
struct WaveLds {
    Survivor slot[kWave];
};
template <>
struct WaveLds<1> {
    Survivor slot[kWave + 1];
    uint32_t list[4][kWave];
};

// Set bits of a wavefront mask below this lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// LDS byte address of a __shared__ object, and a survivor read from one.
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}
typedef float LdsF4 __attribute__((ext_vector_type(4)));
typedef float LdsF2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void lds_survivor(uint32_t addr, float4& a, float4& b, float2& c) {
    const __attribute__((address_space(3))) LdsF4* p4 = (const __attribute__((address_space(3))) LdsF4*)(uintptr_t)addr;
    const LdsF4 va = p4[0], vb = p4[1];
    const LdsF2 vc = *(const __attribute__((address_space(3))) LdsF2*)(uintptr_t)(addr + 32u);
    a = make_float4(va.x, va.y, va.z, va.w);
    b = make_float4(vb.x, vb.y, vb.z, vb.w);
    c = make_float2(vc.x, vc.y);
}

// TW x TW tile; RasterShape gives the wavefronts' pixel blocks. `proj` holds
// one record of `rec` float4s per splat (4: gs3d_projected_splat, 3: a batch
// view's 48-B record); only the first three float4s are read.
//
// Per batch of 64 list entries a wavefront gathers one splat per lane, keeps
// the ones that may touch its block (may_touch) and compacts them, in list
// order, into its own LDS slots; the count is padded to a multiple of 4 with
// zero-opacity entries, which fail alpha >= 1/255 and blend nothing. The
// composite loop reads 4 slots at a time with uniform-address LDS loads (a
// broadcast) and blends branch-free: a splat that fails a rule is blended with
// alpha 0, which leaves C and T bit-identical (T - T * 0 = T,
// C + T * 0 * c = C). The all-saturated exit is checked once per 4 splats.
// One wavefront's pixel block `wave` (0 .. kWaves-1) of `tile`, with that
// wavefront's own LDS slots. This form serves PPL > 1, and one pixel per lane
// in k_rasterize_blocks; k_rasterize at one pixel per lane takes
// raster_block_quarters below.
template <int TW, int PPL>
__device__ __forceinline__ void raster_block_rows(const float4* __restrict__ proj, uint32_t rec,
                                                  const uint32_t* __restrict__ values,
                                                  const uint32_t* __restrict__ tile_offsets, uint32_t tile,
                                                  int wave, uint32_t tiles_x, uint32_t width, uint32_t height,
                                                  uint2* __restrict__ image, Survivor* slots) {
    using Sh = RasterShape<TW, PPL>;
    constexpr float kLog2e = 1.4426950408889634f;

    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t bx0 = tx * TW + (wave % Sh::kBlocksX) * Sh::BW;
    const uint32_t by0 = ty * TW + (wave / Sh::kBlocksX) * Sh::BH;
    const uint32_t px = bx0 + (lane % Sh::BW);
    const uint32_t py0 = by0 + lane / Sh::BW;
    const float pxf = (float)px + 0.5f;

    // pixel-centre rectangle of this wavefront's block
    const float rx0 = (float)bx0 + 0.5f, rx1 = rx0 + (float)(Sh::BW - 1);
    const float ry0 = (float)by0 + 0.5f, ry1 = ry0 + (float)(Sh::BH - 1);

    // A pixel is live while T >= 1e-4: T only changes on a blend and the pixel
    // stops at the first blend that takes T below 1e-4, so that test is the
    // termination flag. Pixels outside the image start at T = 0 (never live).
    float c0[PPL], c1[PPL], c2[PPL], tr[PPL], pyf[PPL];
#pragma unroll
    for (int k = 0; k < PPL; k++) {
        const uint32_t py = py0 + k * Sh::kLaneStride;
        c0[k] = c1[k] = c2[k] = 0.0f;
        tr[k] = (px < width && py < height) ? 1.0f : 0.0f;
        pyf[k] = (float)py + 0.5f;
    }
    constexpr float kTMin = 1.0e-4f;
    auto all_done = [&]() {
        bool l = false;
#pragma unroll
        for (int k = 0; k < PPL; k++) l = l || tr[k] >= kTMin;
        return __ballot(l) == 0ull;
    };
    auto blend = [&](const Survivor& s) {
        const float dx = pxf - s.a.x;
        const float dxa = s.a.w * dx;
        const float dx2 = dx * dx;
#pragma unroll
        for (int k = 0; k < PPL; k++) {
            const float dy = pyf[k] - s.a.y;
            const float u = fmaf(s.b.x, dy, dxa);         // A dx + B dy
            const float p = fmaf(s.b.y * dy, dy, u * dx);  // log2(e) * power
            const float d2 = fmaf(dy, dy, dx2);
            const float alpha = fminf(s.b.z * __builtin_amdgcn_exp2f(p), 0.99f);
            // non-short-circuit: all four tests are evaluated, no branch per splat
            const bool ok = (tr[k] >= kTMin) & (d2 <= s.a.z) & (p <= 0.0f) & (alpha >= 1.0f / 255.0f);
            const float ae = ok ? alpha : 0.0f;
            const float w = tr[k] * ae;
            c0[k] = fmaf(w, s.b.w, c0[k]);
            c1[k] = fmaf(w, s.c.x, c1[k]);
            c2[k] = fmaf(w, s.c.y, c2[k]);
            // T (1 - a) as T - T a: one instruction instead of two, the same
            // value up to rounding; a rejected splat (a = 0) leaves T bit-identical
            tr[k] = tr[k] - w;
        }
    };

    const uint32_t start = tile_offsets[tile];
    const uint32_t end = tile_offsets[tile + 1];

    // software pipeline: batch b+1 is gathered while batch b is composited
    float4 n0 = make_float4(0.f, 0.f, 0.f, 0.f), n1 = n0, n2 = n0;
    if (start < end && start + lane < end) {
        const uint32_t sid = values[start + lane];
        n0 = proj[(size_t)sid * rec + 0];
        n1 = proj[(size_t)sid * rec + 1];
        n2 = proj[(size_t)sid * rec + 2];
    }
    for (uint32_t base = start; base < end; base += kWave) {
        if (all_done()) break;
        const float4 f0 = n0, f1 = n1, f2 = n2;
        const bool have = base + lane < end;
        const uint32_t nidx = base + kWave + lane;
        if (nidx < end) {
            const uint32_t sid = values[nidx];
            n0 = proj[(size_t)sid * rec + 0];
            n1 = proj[(size_t)sid * rec + 1];
            n2 = proj[(size_t)sid * rec + 2];
        }
        const bool hit = have && may_touch(f0, f1, rx0, rx1, ry0, ry1);
        const uint64_t mask = __ballot(hit);
        const int nhit = __popcll(mask);
        const int nhit4 = (nhit + 3) & ~3;
        if (hit) {
            const int slot = (int)__popcll(mask & ((1ull << lane) - 1ull));
            Survivor s;
            s.a = make_float4(f0.x, f0.y, f0.w * f0.w, -0.5f * kLog2e * f1.x);
            s.b = make_float4(-kLog2e * f1.y, -0.5f * kLog2e * f1.z, f1.w, f2.x);
            s.c = make_float2(f2.y, f2.z);
            s.pad = make_float2(0.f, 0.f);
            slots[slot] = s;
        }
        if (lane >= nhit && lane < nhit4) {
            // pad slots [nhit, nhit4), written by lane = slot whether or not that
            // lane hit (a hit lane's own survivor went to a slot below nhit).
            // Zero opacity: alpha = 0 < 1/255, blends nothing.
            Survivor s;
            s.a = make_float4(0.f, 0.f, 0.f, 0.f);
            s.b = s.a;
            s.c = make_float2(0.f, 0.f);
            s.pad = s.c;
            slots[lane] = s;
        }
        __builtin_amdgcn_wave_barrier();
        for (int j = 0; j < nhit4; j += 4) {
            const Survivor s0 = slots[j], s1 = slots[j + 1], s2 = slots[j + 2], s3 = slots[j + 3];
            blend(s0);
            blend(s1);
            blend(s2);
            blend(s3);
            if (all_done()) break;
        }
        __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int k = 0; k < PPL; k++) {
        const uint32_t py = py0 + k * Sh::kLaneStride;
        if (px < width && py < height) {
            const uint32_t h0 = __half_as_ushort(__float2half_rn(c0[k]));
            const uint32_t h1 = __half_as_ushort(__float2half_rn(c1[k]));
            const uint32_t h2 = __half_as_ushort(__float2half_rn(c2[k]));
            const uint32_t h3 = __half_as_ushort(__float2half_rn(1.0f - tr[k]));
            image[(size_t)py * width + px] = make_uint2(h0 | (h1 << 16), h2 | (h3 << 16));
        }
    }
}

// One pixel per lane: the 8 x 8 block of wavefront `wave`, each 16-lane
// quarter of the wavefront owning a 4 x 4 quarter of the block (lane l: pixel
// (l & 3, (l >> 2) & 3) of quarter l >> 4). A batch is culled per quarter
// (may_touch_quarters); the union is compacted into the slots as in
// raster_block_rows, and each quarter gets its own list of the slots that may
// touch it, in list order, padded to the longest list (rounded up to 4) with
// the zero record. The composite loop then reads, per lane, the survivors of
// its own quarter: a splat that reaches one quarter of the block costs one
// step instead of a step for the whole block. Every pixel still blends its
// own tile's list in order, a splat the cull drops could not have passed the
// per-pixel rules, and the zero record blends nothing, so each pixel's result
// is the same as with one list per block.
template <int TW>
__device__ __forceinline__ void raster_block_quarters(const float4* __restrict__ proj, uint32_t rec,
                                                      const uint32_t* __restrict__ values,
                                                      const uint32_t* __restrict__ tile_offsets,
                                                      uint32_t tile, int wave, uint32_t tiles_x,
                                                      uint32_t width, uint32_t height,
                                                      uint2* __restrict__ image, WaveLds<1>* lds) {
    using Sh = RasterShape<TW, 1>;
    static_assert(Sh::BW == 8 && Sh::BH == 8, "one pixel per lane: 8 x 8 blocks");
    constexpr float kLog2e = 1.4426950408889634f;
    constexpr float kTMin = 1.0e-4f;

    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    const int lane = threadIdx.x & (kWave - 1);
    const int quarter = lane >> 4;
    const uint32_t bx0 = tx * TW + (wave % Sh::kBlocksX) * 8u;
    const uint32_t by0 = ty * TW + (wave / Sh::kBlocksX) * 8u;
    const uint32_t px = bx0 + (uint32_t)(quarter & 1) * 4u + (uint32_t)(lane & 3);
    const uint32_t py = by0 + (uint32_t)(quarter >> 1) * 4u + (uint32_t)((lane >> 2) & 3);
    const float pxf = (float)px + 0.5f, pyf = (float)py + 0.5f;
    const float x0 = (float)bx0 + 0.5f, y0 = (float)by0 + 0.5f;  // first pixel centre of the block

    float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
    float tr = (px < width && py < height) ? 1.0f : 0.0f;  // T >= 1e-4: the pixel is live

    if (lane == 0) {
        Survivor z;
        z.a = make_float4(0.f, 0.f, 0.f, 0.f);
        z.b = z.a;
        z.c = make_float2(0.f, 0.f);
        z.pad = z.c;
        lds->slot[kWave] = z;
    }
    const uint32_t zero_addr = lds_addr(&lds->slot[kWave]);
    const uint32_t slot_addr0 = lds_addr(&lds->slot[0]);
    const uint32_t* const my_list = lds->list[quarter];

    auto blend = [&](float4 a, float4 b, float2 c) {
        const float dx = pxf - a.x, dy = pyf - a.y;
        const float u = fmaf(b.x, dy, a.w * dx);       // A dx + B dy
        const float p = fmaf(b.y * dy, dy, u * dx);    // log2(e) * power
        const float d2 = fmaf(dy, dy, dx * dx);
        const float alpha = fminf(b.z * __builtin_amdgcn_exp2f(p), 0.99f);
        const bool ok = (tr >= kTMin) & (d2 <= a.z) & (p <= 0.0f) & (alpha >= 1.0f / 255.0f);
        const float w = tr * (ok ? alpha : 0.0f);
        c0 = fmaf(w, b.w, c0);
        c1 = fmaf(w, c.x, c1);
        c2 = fmaf(w, c.y, c2);
        tr = tr - w;
    };

    const uint32_t start = tile_offsets[tile];
    const uint32_t end = tile_offsets[tile + 1];

    float4 n0 = make_float4(0.f, 0.f, 0.f, 0.f), n1 = n0, n2 = n0;
    if (start < end && start + lane < end) {
        const uint32_t sid = values[start + lane];
        n0 = proj[(size_t)sid * rec + 0];
        n1 = proj[(size_t)sid * rec + 1];
        n2 = proj[(size_t)sid * rec + 2];
    }
    for (uint32_t base = start; base < end; base += kWave) {
        const uint64_t live = __ballot(tr >= kTMin);
        if (live == 0ull) break;
        const float4 f0 = n0, f1 = n1, f2 = n2;
        const bool have = base + lane < end;
        const uint32_t nidx = base + kWave + lane;
        if (nidx < end) {
            const uint32_t sid = values[nidx];
            n0 = proj[(size_t)sid * rec + 0];
            n1 = proj[(size_t)sid * rec + 1];
            n2 = proj[(size_t)sid * rec + 2];
        }
        bool qh[4];
        may_touch_quarters(f0, f1, x0, y0, qh);
        // a quarter whose 16 pixels are all saturated takes no survivors
        uint64_t qm[4];
        uint32_t jmax = 0;
        bool any = false;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            qh[q] = have && qh[q] && ((live >> (16 * q)) & 0xffffull) != 0ull;
            qm[q] = __ballot(qh[q]);
            jmax = max(jmax, ((uint32_t)__popcll(qm[q]) + 3u) & ~3u);
            any = any || qh[q];
        }
        jmax = __builtin_amdgcn_readfirstlane(jmax);
        const uint32_t slot = lanes_below(qm[0] | qm[1] | qm[2] | qm[3]);
        if (any) {
            Survivor s;
            s.a = make_float4(f0.x, f0.y, f0.w * f0.w, -0.5f * kLog2e * f1.x);
            s.b = make_float4(-kLog2e * f1.y, -0.5f * kLog2e * f1.z, f1.w, f2.x);
            s.c = make_float2(f2.y, f2.z);
            s.pad = make_float2(0.f, 0.f);
            lds->slot[slot] = s;
        }
        // every list entry first points at the zero record (64 lanes x 4
        // entries = the four lists), then the survivors' entries overwrite the
        // head of each list; a wavefront's LDS stores complete in order
        reinterpret_cast<uint4*>(&lds->list[0][0])[lane] = make_uint4(zero_addr, zero_addr, zero_addr, zero_addr);
        const uint32_t my_addr = slot_addr0 + slot * (uint32_t)sizeof(Survivor);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (qh[q]) lds->list[q][lanes_below(qm[q])] = my_addr;
        }
        __builtin_amdgcn_wave_barrier();
        // the next 4 entries are read while these 4 blend (index kept inside the list)
        uint4 ad_next = *reinterpret_cast<const uint4*>(&my_list[0]);
        for (uint32_t j = 0; j < jmax; j += 4) {
            const uint4 ad = ad_next;
            ad_next = *reinterpret_cast<const uint4*>(&my_list[min(j + 4u, (uint32_t)kWave - 4u)]);
            float4 a0, b0, a1, b1, a2, b2, a3, b3;
            float2 e0, e1, e2, e3;
            lds_survivor(ad.x, a0, b0, e0);
            lds_survivor(ad.y, a1, b1, e1);
            lds_survivor(ad.z, a2, b2, e2);
            lds_survivor(ad.w, a3, b3, e3);
            blend(a0, b0, e0);
            blend(a1, b1, e1);
            blend(a2, b2, e2);
            blend(a3, b3, e3);
            if (__ballot(tr >= kTMin) == 0ull) break;
        }
        __builtin_amdgcn_wave_barrier();
    }
    // the pixel again from its centre (x + 0.5 truncates to x), so the integer
    // coordinates need not stay live through the loops
    const uint32_t ox = (uint32_t)pxf, oy = (uint32_t)pyf;
    if (ox < width && oy < height) {
        const uint32_t h0 = __half_as_ushort(__float2half_rn(c0));
        const uint32_t h1 = __half_as_ushort(__float2half_rn(c1));
        const uint32_t h2 = __half_as_ushort(__float2half_rn(c2));
        const uint32_t h3 = __half_as_ushort(__float2half_rn(1.0f - tr));
        image[(size_t)oy * width + ox] = make_uint2(h0 | (h1 << 16), h2 | (h3 << 16));
    }
}

template <int TW, int PPL>
__device__ __forceinline__ void raster_block(const float4* __restrict__ proj, uint32_t rec,
                                             const uint32_t* __restrict__ values,
                                             const uint32_t* __restrict__ tile_offsets, uint32_t tile,
                                             int wave, uint32_t tiles_x, uint32_t width, uint32_t height,
                                             uint2* __restrict__ image, WaveLds<PPL>* lds) {
    if constexpr (PPL == 1) {
        raster_block_quarters<TW>(proj, rec, values, tile_offsets, tile, wave, tiles_x, width, height, image,
                                  lds);
    } else {
        raster_block_rows<TW, PPL>(proj, rec, values, tile_offsets, tile, wave, tiles_x, width, height, image,
                                   lds->slot);
    }
}

// Every pixel block of `tile`, one per wavefront of the workgroup.
template <int TW, int PPL>
__device__ __forceinline__ void raster_tile(const float4* __restrict__ proj, uint32_t rec,
                                            const uint32_t* __restrict__ values,
                                            const uint32_t* __restrict__ tile_offsets, uint32_t tile,
                                            uint32_t tiles_x, uint32_t width, uint32_t height,
                                            uint2* __restrict__ image, WaveLds<PPL>* s_lds) {
    const int wave = threadIdx.x / kWave;
    raster_block<TW, PPL>(proj, rec, values, tile_offsets, tile, wave, tiles_x, width, height, image,
                          &s_lds[wave]);
}

// The one-pixel-per-lane kernels are held to 6 wavefronts per SIMD (at most
// 80 VGPRs; the compiler spills a few values that live across the whole
// kernel, outside the composite loop). They are bound by VALU issue and hide
// the LDS latency of the per-lane survivor reads with the extra wavefront:
// 1M splats SH3 1080p, 16 x 16 tiles, 160.8 us at 5 wavefronts, 156.1 at 6.
#define GS3D_RASTER_OCCUPANCY __attribute__((amdgpu_waves_per_eu(PPL == 1 ? 6 : 1)))

// One workgroup per tile (the single-view frame).
template <int TW, int PPL>
__global__ __launch_bounds__(TW * TW / PPL) GS3D_RASTER_OCCUPANCY void k_rasterize(
    const float4* __restrict__ proj, uint32_t rec, const uint32_t* __restrict__ values,
    const uint32_t* __restrict__ tile_offsets, uint32_t tiles_x, uint32_t width, uint32_t height,
    uint2* __restrict__ image) {
    __shared__ WaveLds<PPL> s_lds[RasterShape<TW, PPL>::kWaves];
    raster_tile<TW, PPL>(proj, rec, values, tile_offsets, blockIdx.x, tiles_x, width, height, image, s_lds);
}

// One workgroup per tile, one pixel per lane, one survivor list per 8 x 8
// block (raster_block_rows; the "QUARTER_LISTS" 0 variant). Per-quarter lists
// pay when splats are small against a block; when most splats span all four
// quarters the quarter test is cost without a saving, and this form is
// faster: 1M SH3 1080p (median radius 6 px) 177 us here against 164 with
// quarter lists; 6M SH3 2160p (median 21 px) 499 us against 553.
template <int TW>
__global__ __launch_bounds__(TW * TW) void k_rasterize_blocks(const float4* __restrict__ proj, uint32_t rec,
                                                              const uint32_t* __restrict__ values,
                                                              const uint32_t* __restrict__ tile_offsets,
                                                              uint32_t tiles_x, uint32_t width, uint32_t height,
                                                              uint2* __restrict__ image) {
    __shared__ Survivor s_slot[RasterShape<TW, 1>::kWaves][kWave];
    const int wave = threadIdx.x / kWave;
    raster_block_rows<TW, 1>(proj, rec, values, tile_offsets, blockIdx.x, wave, tiles_x, width, height, image,
                             s_slot[wave]);
}

// The batch's persistent rasterizer: a grid sized to what the GPU holds at
// once, every workgroup taking (view, tile) items from `queue` in order with
// a ticket until the queue is exhausted. Tickets are taken one item ahead
// (double-buffered in LDS) so the atomic's latency hides behind a tile. No
// workgroup waits on another, so any grid size drains the queue.
// The wave-item form below takes pixel blocks instead of tiles: each
// wavefront of a 256-thread workgroup takes its own tickets (item = tile item
// * blocks per tile + block), so no wavefront waits at a barrier for the
// slowest block of its tile.
template <int TW, int PPL>
__global__ __launch_bounds__(TW * TW / PPL) GS3D_RASTER_OCCUPANCY void k_rasterize_queue(
    const RasterView* __restrict__ views, const uint2* __restrict__ queue, uint32_t n_items, uint32_t* ticket) {
    __shared__ WaveLds<PPL> s_lds[RasterShape<TW, PPL>::kWaves];
    __shared__ uint32_t s_next[2];
    if (threadIdx.x == 0) s_next[0] = atomicAdd(ticket, 1u);
    __syncthreads();
    for (uint32_t it = 0;; it++) {
        const uint32_t item = s_next[it & 1u];
        if (item >= n_items) break;  // uniform: every thread read the same word
        if (threadIdx.x == 0) s_next[(it + 1u) & 1u] = atomicAdd(ticket, 1u);
        const uint2 q = queue[item];
        const RasterView& v = views[q.x];
        raster_tile<TW, PPL>(v.proj, 3, v.values, v.tile_offsets, q.y, v.tiles_x, v.width, v.height, v.image,
                             s_lds);
        // the next ticket is visible to all; the slot read at the top of this
        // iteration is the one thread 0 overwrites in the next
        __syncthreads();
    }
}

template <int TW, int PPL>
__global__ __launch_bounds__(256) GS3D_RASTER_OCCUPANCY void k_rasterize_queue_waves(
    const RasterView* __restrict__ views, const uint2* __restrict__ queue, uint32_t n_items, uint32_t* ticket) {
    constexpr uint32_t kBlocks = RasterShape<TW, PPL>::kWaves;
    __shared__ WaveLds<PPL> s_lds[256 / kWave];
    const int lane = threadIdx.x & (kWave - 1);
    WaveLds<PPL>* const lds = &s_lds[threadIdx.x / kWave];
    const uint32_t total = n_items * kBlocks;
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(ticket, 1u);
    uint32_t item = __shfl(t, 0, kWave);
    while (item < total) {  // uniform over the wavefront
        if (lane == 0) t = atomicAdd(ticket, 1u);  // the next item, in flight during this one
        const uint2 q = queue[item / kBlocks];
        const RasterView& v = views[q.x];
        raster_block<TW, PPL>(v.proj, 3, v.values, v.tile_offsets, q.y, (int)(item % kBlocks), v.tiles_x,
                              v.width, v.height, v.image, lds);
        item = __shfl(t, 0, kWave);
    }
}

// ---- the batch's raster queue: (view, tile) items, longest tile lists first ----
//
// A tile's list [off[t], off[t+1]) is empty when off[t+1] <= off[t] (the tiles
// after the last populated one keep zero). Bucket = bit length of the list
// length (0 for empty), placed in descending bucket order; within a bucket the
// order is whatever the atomics give, which changes no pixel: every tile's
// result depends on its own list only.
constexpr uint32_t kQueueBuckets = 33;

__device__ __forceinline__ uint32_t list_bucket(const uint32_t* off, uint32_t t) {
    const uint32_t a = off[t], b = off[t + 1];
    return b > a ? 32u - (uint32_t)__clz(b - a) : 0u;
}

// grid (ceil(max tiles / 256), views): per workgroup, bucket counts in LDS,
// one global add per non-empty bucket.
__global__ __launch_bounds__(256) void k_queue_count(const RasterView* __restrict__ views,
                                                     uint32_t* __restrict__ counts) {
    __shared__ uint32_t s_count[kQueueBuckets];
    if (threadIdx.x < kQueueBuckets) s_count[threadIdx.x] = 0u;
    __syncthreads();
    const RasterView& v = views[blockIdx.y];
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t < v.tiles) atomicAdd(&s_count[list_bucket(v.tile_offsets, t)], 1u);
    __syncthreads();
    if (threadIdx.x < kQueueBuckets && s_count[threadIdx.x]) atomicAdd(&counts[threadIdx.x], s_count[threadIdx.x]);
}

// One wavefront: cursor[b] = items in longer buckets (exclusive scan from the top bucket down).
__global__ void k_queue_place(const uint32_t* __restrict__ counts, uint32_t* __restrict__ cursor) {
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int b = (int)kQueueBuckets - 1; b >= 0; b--) {
            cursor[b] = run;
            run += counts[b];
        }
    }
}

// Same grid as k_queue_count: each workgroup reserves a range per bucket with
// one global add and places its items in it.
__global__ __launch_bounds__(256) void k_queue_scatter(const RasterView* __restrict__ views,
                                                       uint32_t* __restrict__ cursor, uint2* __restrict__ queue) {
    __shared__ uint32_t s_count[kQueueBuckets];
    __shared__ uint32_t s_base[kQueueBuckets];
    if (threadIdx.x < kQueueBuckets) s_count[threadIdx.x] = 0u;
    __syncthreads();
    const RasterView& v = views[blockIdx.y];
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    uint32_t b = 0, rank = 0;
    const bool live = t < v.tiles;
    if (live) {
        b = list_bucket(v.tile_offsets, t);
        rank = atomicAdd(&s_count[b], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kQueueBuckets)
        s_base[threadIdx.x] = s_count[threadIdx.x] ? atomicAdd(&cursor[threadIdx.x], s_count[threadIdx.x]) : 0u;
    __syncthreads();
    if (live) queue[s_base[b] + rank] = make_uint2(blockIdx.y, t);
}

__global__ void k_f16_to_f32(const uint16_t* in, float* out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = __half2float(__ushort_as_half(in[i]));
}

}  // namespace

LaunchShape identify_launch_shape(uint32_t tiles) { return {256u, ((uint64_t)tiles + 1u + 255u) / 256u}; }
LaunchShape raster_launch_shape(uint32_t tile, uint32_t pixels_per_lane, uint32_t tiles) {
    return {tile * tile / (pixels_per_lane ? pixels_per_lane : 1u), tiles};
}

hipError_t launch_identify_tiles(const void* keys, uint32_t key_bytes, const uint32_t* n_dev, uint32_t n_max,
                                 uint32_t* tile_offsets, uint32_t tile_count, uint32_t* k_host, hipStream_t s) {
    const uint32_t threads = tile_count + 1;
    const uint32_t stride = key_bytes == 8 ? 2u : 1u, word = key_bytes == 8 ? 1u : 0u;
    (void)threads;
    const dim3 grid((uint32_t)identify_launch_shape(tile_count).grid);
    if (key_bytes == 2) {
        if (tile_count > 65536u) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_identify<uint16_t>, grid, dim3(256), 0, s, static_cast<const uint16_t*>(keys), 1u, 0u,
                           n_dev, tile_offsets, tile_count, n_max, k_host);
    } else {
        hipLaunchKernelGGL(k_identify<uint32_t>, grid, dim3(256), 0, s, static_cast<const uint32_t*>(keys), stride,
                           word, n_dev, tile_offsets, tile_count, n_max, k_host);
    }
    return hipGetLastError();
}

hipError_t launch_rebuild_keys(const void* tiles, uint32_t tile_bytes, const uint32_t* values,
                               const gs3d_projected_splat* projected, const uint32_t* n_dev, uint32_t n_max,
                               uint64_t* keys, hipStream_t s) {
    if (n_max == 0) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((n_max + 255u) / 256u, 8192u);
    if (tile_bytes == 2)
        hipLaunchKernelGGL(k_rebuild_keys<uint16_t>, dim3(blocks), dim3(256), 0, s,
                           static_cast<const uint16_t*>(tiles), values, projected, n_dev, n_max, keys);
    else if (tile_bytes == 4)
        hipLaunchKernelGGL(k_rebuild_keys<uint32_t>, dim3(blocks), dim3(256), 0, s,
                           static_cast<const uint32_t*>(tiles), values, projected, n_dev, n_max, keys);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <int TW, int PPL>
static void launch_raster_variant(uint32_t tiles, const float4* p, uint32_t rec, const uint32_t* values,
                                  const uint32_t* tile_offsets, uint32_t tiles_x, uint32_t width,
                                  uint32_t height, uint2* img, hipStream_t s) {
    hipLaunchKernelGGL((k_rasterize<TW, PPL>), dim3(tiles), dim3(TW * TW / PPL), 0, s, p, rec, values,
                       tile_offsets, tiles_x, width, height, img);
}

template <int TW>
static void launch_raster_blocks(uint32_t tiles, const float4* p, uint32_t rec, const uint32_t* values,
                                 const uint32_t* tile_offsets, uint32_t tiles_x, uint32_t width, uint32_t height,
                                 uint2* img, hipStream_t s) {
    hipLaunchKernelGGL((k_rasterize_blocks<TW>), dim3(tiles), dim3(TW * TW), 0, s, p, rec, values, tile_offsets,
                       tiles_x, width, height, img);
}

bool raster_variant_exists(uint32_t tile, uint32_t ppl) {
    if (tile != 8 && tile != 16 && tile != 32) return false;
    if (ppl != 1 && ppl != 2 && ppl != 4) return false;
    return tile * tile / ppl >= (uint32_t)kWave;
}

hipError_t launch_rasterize(const gs3d_projected_splat* projected, const uint32_t* values,
                            const uint32_t* tile_offsets, uint32_t tile_x, uint32_t tile_y,
                            uint32_t tiles_x, uint32_t tiles_y, uint32_t width, uint32_t height,
                            uint32_t pixels_per_lane, uint16_t* image_rgba16f, hipStream_t s,
                            bool quarter_lists) {
    return launch_rasterize_records(reinterpret_cast<const float4*>(projected), 4, values, tile_offsets, tile_x,
                                    tile_y, tiles_x, tiles_y, width, height, pixels_per_lane, image_rgba16f, s,
                                    quarter_lists);
}

hipError_t launch_rasterize_records(const float4* p, uint32_t rec, const uint32_t* values,
                                    const uint32_t* tile_offsets, uint32_t tile_x, uint32_t tile_y,
                                    uint32_t tiles_x, uint32_t tiles_y, uint32_t width, uint32_t height,
                                    uint32_t pixels_per_lane, uint16_t* image_rgba16f, hipStream_t s,
                                    bool quarter_lists) {
    const uint32_t tiles = tiles_x * tiles_y;
    if (tiles == 0) return hipSuccess;
    if (tile_x != tile_y || !raster_variant_exists(tile_x, pixels_per_lane)) return hipErrorInvalidValue;
    if (rec != 3 && rec != 4) return hipErrorInvalidValue;
    uint2* img = reinterpret_cast<uint2*>(image_rgba16f);
    if (pixels_per_lane == 1 && !quarter_lists) {
        switch (tile_x) {
            case 8: launch_raster_blocks<8>(tiles, p, rec, values, tile_offsets, tiles_x, width, height, img, s); break;
            case 16: launch_raster_blocks<16>(tiles, p, rec, values, tile_offsets, tiles_x, width, height, img, s); break;
            case 32: launch_raster_blocks<32>(tiles, p, rec, values, tile_offsets, tiles_x, width, height, img, s); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    const int key = tile_x * 10 + pixels_per_lane;
    switch (key) {
        case 81: launch_raster_variant<8, 1>(tiles, p, rec, values, tile_offsets, tiles_x, width, height, img, s); break;
        case 161: launch_raster_variant<16, 1>(tiles, p, rec, values, tile_offsets, tiles_x, width, height, img, s); break;
        case 162: launch_raster_variant<16, 2>(tiles, p, rec, values, tile_offsets, tiles_x, width, height, img, s); break;
        case 164: launch_raster_variant<16, 4>(tiles, p, rec, values, tile_offsets, tiles_x, width, height, img, s); break;
        case 321: launch_raster_variant<32, 1>(tiles, p, rec, values, tile_offsets, tiles_x, width, height, img, s); break;
        case 322: launch_raster_variant<32, 2>(tiles, p, rec, values, tile_offsets, tiles_x, width, height, img, s); break;
        case 324: launch_raster_variant<32, 4>(tiles, p, rec, values, tile_offsets, tiles_x, width, height, img, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_raster_queue_build(const RasterView* views, uint32_t n_views, uint32_t max_tiles,
                                     uint32_t* ctl, uint2* queue, hipStream_t s) {
    if (n_views == 0 || max_tiles == 0) return hipSuccess;
    uint32_t* counts = ctl + kRasterQueueCounts;
    uint32_t* cursor = ctl + kRasterQueueCursor;
    const dim3 grid((max_tiles + 255u) / 256u, n_views);
    hipLaunchKernelGGL(k_queue_count, grid, dim3(256), 0, s, views, counts);
    hipLaunchKernelGGL(k_queue_place, dim3(1), dim3(64), 0, s, counts, cursor);
    hipLaunchKernelGGL(k_queue_scatter, grid, dim3(256), 0, s, views, cursor, queue);
    return hipGetLastError();
}

template <int TW, int PPL>
static hipError_t queue_grid_variant(uint32_t wg_per_cu, bool wave_items, uint32_t* grid) {
    int per_cu = 0, cus = 0, dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = wave_items ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rasterize_queue_waves<TW, PPL>, 256, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rasterize_queue<TW, PPL>,
                                                                      TW * TW / PPL, 0);
    if (e != hipSuccess) return e;
    if (wg_per_cu > 0) per_cu = std::min(per_cu, (int)wg_per_cu);
    *grid = (uint32_t)std::max(1, per_cu * cus);
    return hipSuccess;
}

hipError_t rasterize_queue_grid(uint32_t tile, uint32_t pixels_per_lane, uint32_t wg_per_cu, bool wave_items,
                                uint32_t* grid) {
    if (!raster_variant_exists(tile, pixels_per_lane)) return hipErrorInvalidValue;
    switch (tile * 10 + pixels_per_lane) {
        case 81: return queue_grid_variant<8, 1>(wg_per_cu, wave_items, grid);
        case 161: return queue_grid_variant<16, 1>(wg_per_cu, wave_items, grid);
        case 162: return queue_grid_variant<16, 2>(wg_per_cu, wave_items, grid);
        case 164: return queue_grid_variant<16, 4>(wg_per_cu, wave_items, grid);
        case 321: return queue_grid_variant<32, 1>(wg_per_cu, wave_items, grid);
        case 322: return queue_grid_variant<32, 2>(wg_per_cu, wave_items, grid);
        case 324: return queue_grid_variant<32, 4>(wg_per_cu, wave_items, grid);
        default: return hipErrorInvalidValue;
    }
}

template <int TW, int PPL>
static void launch_queue_variant(uint32_t grid, bool wave_items, const RasterView* views, const uint2* queue,
                                 uint32_t n_items, uint32_t* ticket, hipStream_t s) {
    if (wave_items)
        hipLaunchKernelGGL((k_rasterize_queue_waves<TW, PPL>), dim3(grid), dim3(256), 0, s, views, queue, n_items,
                           ticket);
    else
        hipLaunchKernelGGL((k_rasterize_queue<TW, PPL>), dim3(grid), dim3(TW * TW / PPL), 0, s, views, queue,
                           n_items, ticket);
}

hipError_t launch_rasterize_queue(const RasterView* views, const uint2* queue, uint32_t n_items,
                                  uint32_t* ctl, uint32_t tile, uint32_t pixels_per_lane, uint32_t grid,
                                  bool wave_items, hipStream_t s) {
    if (n_items == 0) return hipSuccess;
    if (!raster_variant_exists(tile, pixels_per_lane) || grid == 0) return hipErrorInvalidValue;
    if (!wave_items) grid = std::min(grid, n_items);
    uint32_t* ticket = ctl + kRasterQueueTicket;
    switch (tile * 10 + pixels_per_lane) {
        case 81: launch_queue_variant<8, 1>(grid, wave_items, views, queue, n_items, ticket, s); break;
        case 161: launch_queue_variant<16, 1>(grid, wave_items, views, queue, n_items, ticket, s); break;
        case 162: launch_queue_variant<16, 2>(grid, wave_items, views, queue, n_items, ticket, s); break;
        case 164: launch_queue_variant<16, 4>(grid, wave_items, views, queue, n_items, ticket, s); break;
        case 321: launch_queue_variant<32, 1>(grid, wave_items, views, queue, n_items, ticket, s); break;
        case 322: launch_queue_variant<32, 2>(grid, wave_items, views, queue, n_items, ticket, s); break;
        case 324: launch_queue_variant<32, 4>(grid, wave_items, views, queue, n_items, ticket, s); break;
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
