// runtime.cpp — context, named device buffers, stage profiler and the C ABI.
//
// A gs3d_ctx is the native counterpart of the reference renderer's resource
// services: BufferManager's named buffers (src/BufferManager.ts:32-154; here a
// name -> device allocation map that only grows, freeing the old block), the
// Profiler's per-pass timestamps (src/Profiler.ts:135-228; here hipEvent pairs
// per stage on the context's stream) and GaussianSplatRenderer.rebin
// (src/renderers/GaussianSplatRenderer.ts:1104-1266), which queues the whole
// pipeline on one HIP stream with no host round trip inside the frame.
//
// Key count K lives on the device (ref_counter); the sort, identify and
// raster launches read it there and size their grids from the key-buffer
// capacity. The identify kernel also stores K into host-mapped memory; when a
// caller waits for results (sync, read) and K exceeded the capacity (the emit
// kernel counts but does not store past it), the buffers grow and the last
// frame is re-run before any result is handed out. A frame queues no copies
// and no events unless it is profiled.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "gs3d_internal.h"

using namespace gs3d;

namespace {

constexpr int kCtrRefCounter = 32;
constexpr int kCtrSplatCount = 33;  // N, for kernels that read their element count on the device
constexpr int kCtrRefCounterId = 34;  // K of the id-order scan (reference-layout offsets)
constexpr int kCtrIdScanTicket = 35;
constexpr int kCtrHist = 64;  // [kMaxPasses][256] digit histograms follow the counters
constexpr size_t kCtrTotalWords = kCtrHist + kMaxPasses * 256;

struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;   // bytes allocated
    size_t size = 0;  // bytes in use (the BufferManager-visible size)
};

const char* kStageLabels[GS3D_STAGE__COUNT] = {
    "preprocess-splats", "radix-sort-splats", "prefix-scan", "emit-tile-refs",
    "radix-sort", "identify-tile-ranges", "rasterize-tiles", "rebin-total", "composite-splats"};
// stages a rebin brackets (the composite is timed by gs3d_composite)
constexpr uint32_t kRebinStageMask = (1u << GS3D_STAGE_COMPOSITE) - 1u;

}  // namespace

struct gs3d_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::unordered_map<std::string, DevBuf> bufs;
    uint32_t limits[6] = {0, 0, 0, 0, 0, 0};

    // scene
    uint32_t n = 0;
    int sh_degree = 0;

    // renderer state
    uint32_t tile_x = 16, tile_y = 16;
    uint32_t width = 0, height = 0, tiles_x = 0, tiles_y = 0;
    uint32_t digit_bits = 8;
    uint32_t raster_ppl = 1;
    uint32_t raster_quarters = 2;  // one pixel per lane: survivor lists per 4x4 quarter (1) or per 8x8 block
                                   // (0), or (2) chosen for each frame from the previous frame's
                                   // references per splat (quarter_lists_auto)
    bool auto_quarters = true;     // the choice for the next frame in mode 2
    uint32_t batch_raster_groups = 1;     // persistent raster launches per batch (view groups); 0: one
                                          // raster launch per view on its branch stream
    uint32_t batch_raster_wg_per_cu = 0;  // their workgroups per CU (0: as many as fit)
    uint32_t batch_raster_wave_items = 0; // 0: a workgroup takes a tile per ticket, 1: a wavefront a pixel block
    uint32_t sort_key_tile = 0;    // the frame's tile-id passes (and stage sorts of 64-bit keys): 0, 8192, 16384
    uint32_t sort_key_tile32 = 0;  // 32-bit key passes: 0 = 8192, or 4096
    uint64_t key_capacity = 0;
    uint32_t n_passes = 0;
    uint32_t shifts[kMaxPasses] = {};
    bool result_in_b = false;
    uint32_t frame_tile_bytes = 4;  // bytes per tile id in the last frame's tile-id sort (2 or 4)
    bool keep_unsorted = false;
    uint32_t epoch = 0;
    uint32_t debug_mode = 0, debug_ref = 256;  // setDebugMode, debugRef
    size_t framebuffer_px = 0;                 // pixels the resident framebuffer holds (0: none yet)
    size_t debug_target_px = 0;                // pixels of the last heat-map target

    // last frame (for re-runs after a capacity miss)
    bool have_frame = false;
    uint64_t splat_gen = 0;  // counts uploads: what a batch's outputs were made from
    gs3d_camera last_cam{};
    uint32_t last_w = 0, last_h = 0, last_flags = 0;
    uint32_t* pinned = nullptr;  // host-mapped; [0] = K of the last frame, written by identify
    bool k_pending = false;
    uint64_t K = 0;

    // profiler: a ring of event pairs, one slot per profiled frame; a slot is
    // folded into the running sums before reuse and when averages are read.
    static constexpr int kRing = 64;
    hipEvent_t ring[kRing][GS3D_STAGE__COUNT][2] = {};
    bool ring_unread[kRing] = {};
    uint32_t ring_mask[kRing] = {};
    uint32_t profile_mask = (1u << GS3D_STAGE__COUNT) - 1u;
    uint64_t stage_count[GS3D_STAGE__COUNT] = {};
    uint64_t frame_counter = 0;
    int cur_slot = 0;
    bool profiled = false;
    bool last_profiled = false;
    double stage_sum_us[GS3D_STAGE__COUNT] = {};
    uint64_t stage_frames = 0;
    // the display pass, timed by gs3d_composite after a profiled rebin
    hipEvent_t composite_ev[2] = {};
    bool composite_pending = false;
    float composite_last_us = 0.f;

    struct BatchState* batch = nullptr;  // gs3d_rebin_batch state, created on first use

    // learned colour
    bool learned_init = false;
    bool learned_enabled = false;
    bool learned_force = false;  // gs3d_learned_validate's learned render
    uint32_t learned_n = 0;      // splats the latents belong to
    uint32_t learned_step = 0;   // Adam steps taken
};

namespace {

int fail(gs3d_ctx* c, int code, const char* fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

#define HIP_TRY(ctx, expr)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail((ctx), e_ == hipErrorOutOfMemory ? GS3D_E_OOM : GS3D_E_HIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                     \
    } while (0)

#define GS3D_TRY(expr)          \
    do {                        \
        int r_ = (expr);        \
        if (r_ != GS3D_OK) return r_; \
    } while (0)

// Lookback epochs are process-wide, so a status block handed from one
// context to another can never hold a word of a live epoch.
std::atomic<uint32_t> g_epoch{0};

uint32_t next_epoch(gs3d_ctx* c) {
    uint32_t e;
    do {
        e = (g_epoch.fetch_add(1u) + 1u) & 0x3FFFFFFFu;
    } while (e == 0);
    c->epoch = e;
    return e;
}

// BufferManager.resize semantics with reuse: grow-only, old block freed.
int ensure(gs3d_ctx* c, const char* name, size_t bytes, void** out) {
    DevBuf& b = c->bufs[name];
    if (bytes == 0) bytes = 4;
    if (b.cap < bytes) {
        if (b.ptr) {
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            HIP_TRY(c, hipFree(b.ptr));
            b.ptr = nullptr;
            b.cap = 0;
        }
        hipError_t e = hipMalloc(&b.ptr, bytes);
        if (e != hipSuccess)
            return fail(c, GS3D_E_OOM, "allocating %zu bytes for '%s': %s", bytes, name,
                        hipGetErrorString(e));
        b.cap = bytes;
    }
    b.size = bytes;
    *out = b.ptr;
    return GS3D_OK;
}

template <class T>
int ensure_t(gs3d_ctx* c, const char* name, size_t count, T** out) {
    void* p = nullptr;
    GS3D_TRY(ensure(c, name, count * sizeof(T), &p));
    *out = static_cast<T*>(p);
    return GS3D_OK;
}

// Lookback status words: zeroed whenever the block is (re)allocated, so
// nothing left in recycled device memory can read as a published status.
int ensure_status(gs3d_ctx* c, const char* name, size_t count, uint64_t** out) {
    const size_t before = c->bufs[name].cap;
    GS3D_TRY(ensure_t(c, name, count, out));
    if (c->bufs[name].cap != before) {
        HIP_TRY(c, hipMemsetAsync(*out, 0, c->bufs[name].cap, c->stream));
    }
    return GS3D_OK;
}

template <class T>
T* buf(gs3d_ctx* c, const char* name) {
    auto it = c->bufs.find(name);
    return it == c->bufs.end() ? nullptr : static_cast<T*>(it->second.ptr);
}

float half_to_float_host(uint16_t h) {
    const uint32_t sign = (uint32_t)(h >> 15) << 31, exp = (h >> 10) & 0x1Fu, man = h & 0x3FFu;
    uint32_t bits;
    if (exp == 0) {
        if (man == 0) {
            bits = sign;
        } else {  // subnormal: renormalise
            int e = -1;
            uint32_t m = man;
            do {
                e++;
                m <<= 1;
            } while (!(m & 0x400u));
            bits = sign | ((uint32_t)(127 - 15 - e) << 23) | ((m & 0x3FFu) << 13);
        }
    } else if (exp == 31) {
        bits = sign | 0x7F800000u | (man << 13);
    } else {
        bits = sign | ((exp + 112u) << 23) | (man << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

uint32_t bit_length(uint32_t v) {
    uint32_t b = 0;
    while (v) {
        b++;
        v >>= 1;
    }
    return b;
}

// Radix passes for keys whose tile id is < tile_count: the 32 depth bits
// always, the tile bits up to bit_length(tile_count - 1). Passes over digits
// that are zero for every key are result-identical no-ops and are not run.
void plan_passes(gs3d_ctx* c, uint32_t tile_count, uint32_t bits) {
    const uint32_t tile_bits = tile_count > 1 ? bit_length(tile_count - 1) : 0;
    c->n_passes = 0;
    for (uint32_t s = 0; s < 32u + tile_bits; s += bits) c->shifts[c->n_passes++] = s;
}

// List form of the next frame in QUARTER_LISTS mode 2. Per-quarter lists pay
// when splats are small against an 8 x 8 block; when most splats span the
// block, the quarter test costs VALU and saves no blend steps (1M SH3 1080p:
// 163 us with quarter lists, 178 with one list per block; 6M SH3 2160p: 550
// against 503). References per splat, scaled to 16 x 16 tiles, measure the
// splats' size: 3.6 and 19.8 on those two frames; above 8, one list per block.
// Both forms give the same image bit for bit, so only the time depends on
// the choice.
bool quarter_lists_auto(uint64_t K, uint32_t n, uint32_t tile_x, uint32_t tile_y) {
    return (double)K * tile_x * tile_y <= 8.0 * 256.0 * (double)n;
}

bool frame_quarter_lists(const gs3d_ctx* c) {
    return c->raster_quarters == 2 ? c->auto_quarters : c->raster_quarters != 0;
}

void record(gs3d_ctx* c, int stage, int which) {
    if (c->profiled && (c->profile_mask >> stage & 1u))
        (void)hipEventRecord(c->ring[c->cur_slot][stage][which], c->stream);
}

// Add a completed profiled frame's stage times to the running sums.
int fold_slot(gs3d_ctx* c, int slot) {
    if (!c->ring_unread[slot]) return GS3D_OK;
    for (int s = 0; s < GS3D_STAGE__COUNT; s++) {
        if (!(c->ring_mask[slot] >> s & 1u)) continue;
        HIP_TRY(c, hipEventSynchronize(c->ring[slot][s][1]));
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, c->ring[slot][s][0], c->ring[slot][s][1]));
        c->stage_sum_us[s] += (double)ms * 1000.0;
        c->stage_count[s]++;
    }
    c->stage_frames++;
    c->ring_unread[slot] = false;
    return GS3D_OK;
}

// Add the last timed composite to the running sums.
int fold_composite(gs3d_ctx* c) {
    if (!c->composite_pending) return GS3D_OK;
    HIP_TRY(c, hipEventSynchronize(c->composite_ev[1]));
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->composite_ev[0], c->composite_ev[1]));
    c->composite_last_us = ms * 1000.0f;
    c->stage_sum_us[GS3D_STAGE_COMPOSITE] += (double)ms * 1000.0;
    c->stage_count[GS3D_STAGE_COMPOSITE]++;
    c->composite_pending = false;
    return GS3D_OK;
}

int begin_profiled_frame(gs3d_ctx* c) {
    if (!c->profiled) return GS3D_OK;
    c->cur_slot = (int)(c->frame_counter++ % gs3d_ctx::kRing);
    GS3D_TRY(fold_slot(c, c->cur_slot));  // 64 frames old: long finished
    return GS3D_OK;
}

int set_device(gs3d_ctx* c) {
    HIP_TRY(c, hipSetDevice(c->device));
    return GS3D_OK;
}

// Run LSD passes shifts[0..passes) over n = *n_dev pairs (grid sized for
// n_max), ping-ponging a <-> b. hist: [passes][radix] digit counts of the
// input; ticket: one counter per pass; first_iota: pass 0 takes the
// positions as values; aux_in/aux_out (or null): the last pass also writes
// aux_out[p] = aux_in[value at p]. Sets *in_b when the result ends in b.
// Keys per workgroup of the tile-id passes (32-bit keys): 16,384 unless set.
uint32_t tile_key_tile(const gs3d_ctx* c) { return c->sort_key_tile ? c->sort_key_tile : 16384u; }

// Key buffers hold at most 2^32 - 1 entries: K is a 32-bit count (the
// reference's u32 ref_counter) and the kernels take the capacity as a 32-bit
// bound, so a larger capacity would wrap there.
uint64_t key_capacity_for(uint64_t want) { return std::min<uint64_t>(want, 0xFFFFFFFFull); }

int run_passes(gs3d_ctx* c, void* ka, uint32_t* va, void* kb, uint32_t* vb, uint32_t key_bytes,
               bool first_iota, const uint32_t* n_dev, uint64_t n_max, const uint32_t* shifts,
               uint32_t passes, const uint32_t* hist, uint32_t* tickets, bool* in_b,
               const uint32_t* aux_in = nullptr, uint32_t* aux_out = nullptr, uint32_t key_bits = 64,
               uint32_t key_tile = 0, uint32_t* carry_a = nullptr, uint32_t* carry_b = nullptr) {
    // carry_a / carry_b (or null): a 32-bit array that every pass moves with the
    // keys, starting in carry_a and ending beside the keys (in carry_b exactly
    // when *in_b is set)
    uint64_t* radix_status = nullptr;
    const uint32_t radix = 1u << c->digit_bits;
    GS3D_TRY(ensure_status(c, "radix_status",
                           (size_t)radix_status_tiles((uint32_t)n_max) * radix, &radix_status));
    for (uint32_t p = 0; p < passes; p++) {
        const bool even = (p % 2) == 0;
        const uint32_t* ain = p + 1 == passes ? aux_in : nullptr;
        uint32_t* aout = p + 1 == passes ? aux_out : nullptr;
        if (carry_a) {
            ain = even ? carry_a : carry_b;
            aout = even ? carry_b : carry_a;
        }
        HIP_TRY(c, launch_radix_pass(even ? ka : kb, even ? va : vb, even ? kb : ka, even ? vb : va,
                                     key_bytes, first_iota && p == 0, n_dev, (uint32_t)n_max,
                                     hist + p * radix, radix_status, tickets + p, shifts[p],
                                     c->digit_bits, next_epoch(c), ain, aout, c->stream, nullptr, key_tile,
                                     pass_match_bits(shifts[p], c->digit_bits, key_bits), carry_a != nullptr));
    }
    *in_b = (passes % 2) == 1;
    return GS3D_OK;
}

// Reference-layout offsets: the exclusive scan of counts in splat-id order
// (prefix_scan_*.wgsl). The frame itself scans in depth order; this one runs
// on demand for readback and for GS3D_REBIN_KEEP_UNSORTED.
int enqueue_id_order_scan(gs3d_ctx* c, uint32_t** out) {
    uint32_t *counts = buf<uint32_t>(c, "splat_ref_counts"), *counters = buf<uint32_t>(c, "counters");
    uint32_t* offs;
    uint64_t* st;
    GS3D_TRY(ensure_t(c, "splat_ref_offsets", (size_t)c->n + 1, &offs));
    GS3D_TRY(ensure_status(c, "scan_status_id", (size_t)c->n / kScanTile + 2, &st));
    HIP_TRY(c, launch_fill_u32(counters + kCtrIdScanTicket, 0u, 1, c->stream));
    HIP_TRY(c, launch_scan(counts, nullptr, offs, counters + kCtrRefCounterId, st,
                           counters + kCtrIdScanTicket, c->n, next_epoch(c), c->stream));
    *out = offs;
    return GS3D_OK;
}

// One frame (rebin, GaussianSplatRenderer.ts:1104-1266), all on the context's
// stream, no host round trip:
//   preprocess -> depth keys of the N splats
//   LSD over the N depth keys (digits of bits 0..31), values = splat ids
//   scan of the tile counts in depth order, counted from the tile rectangles
//   (-> K, one record per depth position)
//   emit (tile, depth) keys in depth order from the records (+ tile-digit
//   histograms)
//   LSD over the K keys, tile digits only (bits 32..)
//   identify tile ranges, rasterize
// Sorting the 32 depth bits on the splats instead of on the keys is
// result-identical to the full-width LSD sort of the keys: every key of a
// splat carries the splat's depth, and emitting splats in stable depth order,
// each splat's tiles in row-major order, yields exactly the key stream a
// stable sort of the id-order emit by those 32 bits would; the tile-digit
// passes then see the same input.
int enqueue_rebin(gs3d_ctx* c, const gs3d_camera* cam, uint32_t W, uint32_t H, uint32_t flags) {
    const uint32_t n = c->n;
    c->width = W;
    c->height = H;
    c->tiles_x = (W + c->tile_x - 1) / c->tile_x;
    c->tiles_y = (H + c->tile_y - 1) / c->tile_y;
    const uint32_t T = c->tiles_x * c->tiles_y;
    c->profiled = (flags & GS3D_REBIN_PROFILE) != 0;
    c->keep_unsorted = (flags & GS3D_REBIN_KEEP_UNSORTED) != 0;
    GS3D_TRY(begin_profiled_frame(c));
    if (c->key_capacity == 0) c->key_capacity = key_capacity_for(std::max<uint64_t>(8ull * n, 1u << 16));
    const uint64_t cap = c->key_capacity;
    const uint32_t bits = c->digit_bits;
    const uint32_t radix = 1u << bits;

    gs3d_projected_splat* projected;
    uint32_t *counts, *counters, *tile_offsets, *va, *vb, *dka, *dkb, *ida, *idb;
    uint64_t *scan_status, *ka, *kb;
    uint4* recs;
    uint16_t* image;
    GS3D_TRY(ensure_t(c, "projected_splats", n, &projected));
    GS3D_TRY(ensure_t(c, "splat_ref_counts", n, &counts));
    // Tile rectangles: packed 8 bits per coordinate when both tile counts fit,
    // so the last depth pass carries them into depth order and the scan reads
    // them in order; otherwise 8 B per splat, gathered by the scan.
    const bool rects_packed = c->tiles_x <= 256u && c->tiles_y <= 256u;
    uint2* rects = nullptr;
    uint32_t *rects8 = nullptr, *rects8_sorted = nullptr;
    if (rects_packed) {
        GS3D_TRY(ensure_t(c, "splat_tile_rects8_a", n, &rects8));
        GS3D_TRY(ensure_t(c, "splat_tile_rects8_b", n, &rects8_sorted));
    } else {
        GS3D_TRY(ensure_t(c, "splat_tile_rects", n, &rects));
    }
    GS3D_TRY(ensure_t(c, "counters", kCtrTotalWords, &counters));
    GS3D_TRY(ensure_status(c, "scan_status", (size_t)n / kScanTile + 2, &scan_status));
    GS3D_TRY(ensure_t(c, "splat_emit_records", (size_t)n + 1, &recs));
    GS3D_TRY(ensure_t(c, "depth_keys_a", n, &dka));
    GS3D_TRY(ensure_t(c, "depth_keys_b", n, &dkb));
    GS3D_TRY(ensure_t(c, "splat_order_a", n, &ida));
    GS3D_TRY(ensure_t(c, "splat_order_b", n, &idb));
    GS3D_TRY(ensure_t(c, "sort_keys_a", cap, &ka));
    GS3D_TRY(ensure_t(c, "sort_keys_b", cap, &kb));
    GS3D_TRY(ensure_t(c, "sort_values_a", cap, &va));
    GS3D_TRY(ensure_t(c, "sort_values_b", cap, &vb));
    GS3D_TRY(ensure_t(c, "tile_offsets", (size_t)T + 1, &tile_offsets));
    GS3D_TRY(ensure_t(c, "splat_target", (size_t)W * H * 4, &image));
    plan_passes(c, T, bits);
    const uint32_t depth_passes = 32u / bits;
    const uint32_t tile_passes = c->n_passes - depth_passes;
    uint32_t* ref_counter = counters + kCtrRefCounter;
    uint32_t* hist_depth = counters + kCtrHist;
    uint32_t* hist_tile = counters + kCtrHist + depth_passes * radix;
    uint32_t* hist_partial;
    GS3D_TRY(ensure_t(c, "radix_hist_partial",
                      (size_t)std::max(emit_grid(n), key_hist_rows(n)) * kMaxPasses * radix, &hist_partial));

    record(c, GS3D_STAGE_TOTAL, 0);
    record(c, GS3D_STAGE_PREPROCESS, 0);
    PreprocessArgs pa{};
    pa.counters = counters;  // zeroed by the preprocess kernel itself, N stored
    pa.counter_words = (uint32_t)kCtrTotalWords;
    pa.count_slot = kCtrSplatCount;
    pa.cam = *cam;
    pa.pos = buf<float4>(c, "splat_positions");
    pa.scale = buf<float4>(c, "splat_scales");
    pa.rot = buf<float4>(c, "splat_rotations");
    pa.color = buf<float4>(c, "splat_colors");
    pa.sh = c->sh_degree > 0 ? buf<float>(c, "splat_sh") : nullptr;
    pa.out = projected;
    pa.counts = counts;
    pa.depth_keys = dka;
    pa.rects = rects;
    pa.rects8 = rects8;
    pa.n = n;
    pa.sh_degree = c->sh_degree;
    pa.tiles_x = c->tiles_x;
    pa.tiles_y = c->tiles_y;
    pa.inv_tile_x = (float)(1.0 / (double)c->tile_x);
    pa.inv_tile_y = (float)(1.0 / (double)c->tile_y);
    HIP_TRY(c, launch_preprocess(pa, c->stream));
    record(c, GS3D_STAGE_PREPROCESS, 1);

    // depth order of the splats (stable: ties keep splat-id order)
    record(c, GS3D_STAGE_DEPTH_SORT, 0);
    HIP_TRY(c, launch_key_histogram(dka, 4, counters + kCtrSplatCount, n, hist_depth, hist_partial,
                                    bits, depth_passes, 0, c->stream));
    bool order_in_b = false;
    // pass 0 takes the positions as values (ida is not read); odd passes write ida
    // the packed rectangles move with the depth keys (rects8 <-> rects8_sorted)
    GS3D_TRY(run_passes(c, dka, ida, dkb, idb, 4, true, counters + kCtrSplatCount, n, c->shifts,
                        depth_passes, hist_depth, counters + kCtrSortTicket, &order_in_b, nullptr, nullptr, 32,
                        c->sort_key_tile32, rects8, rects8_sorted));
    const uint32_t* order = order_in_b ? idb : ida;
    const uint32_t* depth_sorted = order_in_b ? dkb : dka;
    record(c, GS3D_STAGE_DEPTH_SORT, 1);

    // scan of the tile counts in depth order, counted from the tile
    // rectangles; it leaves one record per depth position for emit
    record(c, GS3D_STAGE_SCAN, 0);
    ScanRectsArgs sa{};
    sa.order = order;
    sa.depth_keys = depth_sorted;
    sa.rects = rects;
    sa.rects8 = rects_packed ? (order_in_b ? rects8_sorted : rects8) : nullptr;
    sa.recs = recs;
    sa.ref_counter = ref_counter;
    sa.status = scan_status;
    sa.ticket = counters + kCtrScanTicket;
    sa.n = n;
    sa.tiles_x = c->tiles_x;
    sa.epoch = next_epoch(c);
    HIP_TRY(c, launch_scan_rects(sa, c->stream));
    record(c, GS3D_STAGE_SCAN, 1);

    record(c, GS3D_STAGE_EMIT, 0);
    EmitArgs ea{};
    ea.projected = projected;
    ea.order = order;
    ea.recs = recs;
    // tile ids; the sort_keys buffers hold them as u16 up to 65,536 tiles, else u32
    const uint32_t tb = tile_id_bytes(T);
    c->frame_tile_bytes = tb;
    ea.tiles = tb == 4 ? reinterpret_cast<uint32_t*>(ka) : nullptr;
    ea.tiles16 = tb == 2 ? reinterpret_cast<uint16_t*>(ka) : nullptr;
    ea.values = va;
    ea.hist = tile_passes ? hist_tile : nullptr;
    ea.hist_partial = tile_passes ? hist_partial : nullptr;
    ea.n = n;
    ea.tiles_x = c->tiles_x;
    ea.capacity = cap;
    ea.n_passes = tile_passes;
    ea.digit_bits = bits;
    ea.hist_shift0 = 32;
    HIP_TRY(c, launch_emit(ea, c->stream));
    record(c, GS3D_STAGE_EMIT, 1);

    record(c, GS3D_STAGE_SORT, 0);
    // the tile digits of the keys are the digits of the 32-bit tile ids, shifted by 32
    uint32_t tile_shifts[kMaxPasses];
    for (uint32_t p = 0; p < tile_passes; p++) tile_shifts[p] = c->shifts[depth_passes + p] - 32u;
    GS3D_TRY(run_passes(c, ka, va, kb, vb, tb, false, ref_counter, cap, tile_shifts, tile_passes, hist_tile,
                        counters + kCtrSortTicket + depth_passes, &c->result_in_b, nullptr, nullptr,
                        T > 1 ? bit_length(T - 1) : 0u, tile_key_tile(c)));
    record(c, GS3D_STAGE_SORT, 1);
    const void* sorted_tiles = c->result_in_b ? kb : ka;
    const uint32_t* sorted_vals = c->result_in_b ? vb : va;

    record(c, GS3D_STAGE_IDENTIFY, 0);
    HIP_TRY(c, launch_identify_tiles(sorted_tiles, tb, ref_counter, (uint32_t)cap, tile_offsets, T,
                                     c->pinned, c->stream));
    record(c, GS3D_STAGE_IDENTIFY, 1);

    record(c, GS3D_STAGE_RASTER, 0);
    HIP_TRY(c, launch_rasterize(projected, sorted_vals, tile_offsets, c->tile_x, c->tile_y,
                                c->tiles_x, c->tiles_y, W, H, c->raster_ppl, image, c->stream,
                                frame_quarter_lists(c)));
    record(c, GS3D_STAGE_RASTER, 1);
    record(c, GS3D_STAGE_TOTAL, 1);

    if (c->keep_unsorted) {
        // the reference's emit order (splat ids ascending) for parity checks
        uint32_t* id_offsets;
        GS3D_TRY(enqueue_id_order_scan(c, &id_offsets));
        uint64_t* uk;
        uint32_t* uv;
        GS3D_TRY(ensure_t(c, "unsorted_keys", cap, &uk));
        GS3D_TRY(ensure_t(c, "unsorted_values", cap, &uv));
        EmitArgs eu{};
        eu.projected = projected;
        eu.order = nullptr;
        eu.offsets = id_offsets;
        eu.keys = uk;
        eu.values = uv;
        eu.n = n;
        eu.tiles_x = c->tiles_x;
        eu.capacity = cap;
        eu.digit_bits = bits;
        HIP_TRY(c, launch_emit(eu, c->stream));
    }

    c->k_pending = true;
    if (c->profiled) {
        c->ring_unread[c->cur_slot] = true;
        c->ring_mask[c->cur_slot] = c->profile_mask & kRebinStageMask;
    }
    c->last_profiled = c->profiled;
    return GS3D_OK;
}

// Consume the last frame's K (blocking: the stream is drained first). After a
// capacity miss, grow the key buffers and re-run that frame so its results
// are whole.
int settle(gs3d_ctx* c) {
    for (int attempt = 0; attempt < 3; attempt++) {
        if (!c->k_pending) return GS3D_OK;
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        c->k_pending = false;
        c->K = c->pinned[0];
        c->auto_quarters = quarter_lists_auto(c->K, c->n, c->tile_x, c->tile_y);
        if (c->K <= c->key_capacity) return GS3D_OK;
        c->key_capacity = key_capacity_for(c->K + c->K / 4 + kSortTile);  // the next ensure() reallocates
        GS3D_TRY(enqueue_rebin(c, &c->last_cam, c->last_w, c->last_h, c->last_flags));
    }
    return fail(c, GS3D_E_STATE, "key buffer did not settle after re-runs");
}

}  // namespace

// ---- batched views ------------------------------------------------------------
//
// gs3d_rebin_batch renders many views of the resident splats. One launch of
// k_preprocess_views projects every splat into every view (each splat's
// attributes and SH coefficients are read once per 64 views). Then each view
// runs the frame's chain (depth LSD, scan, emit, tile LSD, identify,
// rasterize) on one of kBatchBranches streams: views are dealt to branches
// longest-first by pixel count, so the latency-bound tails of one view's
// kernels overlap with another view's work. Per-view outputs (projected
// splats, tile ranges, image, K) persist; sort scratch is per branch. The
// whole sequence is captured once into a hipGraph and replayed; cameras live
// in device memory, so a new set of cameras replays the same graph.

constexpr int kBatchBranches = 4;
// at most this many view groups, each rasterized by one persistent launch
constexpr uint32_t kMaxRasterGroups = 16;
// a batch view's record per splat: the first three float4s of a gs3d_projected_splat
constexpr size_t kBatchRecordBytes = 48;
// Per-view stride (elements) of the batch's depth-key and tile-rectangle
// arrays: a multiple of 4, so every view's slice (and the second packed
// rectangle array inside it) starts 16-B aligned for the scan's uint4 loads.
inline size_t batch_view_stride(uint32_t n) { return ((size_t)n + 3u) & ~size_t(3); }

struct BatchView {
    gs3d_camera cam{};
    uint32_t W = 0, H = 0, tiles_x = 0, tiles_y = 0;
    size_t img_off = 0, to_off = 0;  // offsets into the batch image / tile-range arrays (elements)
    size_t val_off = 0;              // offset of the view's sorted splat ids (elements)
    int branch = 0;
};

struct BatchState {
    std::vector<BatchView> views;
    // what the captured graph was built for
    std::vector<uint64_t> shape;
    uint32_t n = 0, tile = 0, bits = 0, ppl = 0;
    int deg = -1;
    uint64_t cap[kBatchBranches] = {};
    bool captured = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t streams[kBatchBranches] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kBatchBranches] = {};
    hipStream_t raster_stream = nullptr;  // the persistent rasterizer's launches
    hipEvent_t group_done[kMaxRasterGroups][kBatchBranches] = {};
    hipEvent_t raster_join = nullptr;
    uint32_t groups = 1;  // view groups of the captured graph
    uint32_t wg_per_cu = 0;  // and the rasterizer's workgroups per CU
    uint32_t wave_items = 0;  // and its work-item size
    uint32_t* pinned_k = nullptr;  // host-mapped, one K per view (written by identify)
    uint32_t pinned_words = 0;
    std::vector<uint64_t> K;
    bool have = false;
    uint64_t replays = 0;
    bool learned = false;  // the captured graph computes learned colours
    uint64_t splat_gen = 0;  // the splats the last batch projected
    uint32_t epoch_bumps = 0;  // launches since the replay base was last zeroed
    size_t val_total = 0;      // entries of batch_sorted_values (each view's branch capacity)
    uint32_t n_items = 0;      // raster queue items: the tiles of every view
    uint32_t max_tiles = 0;    // tiles of the largest view
    uint32_t raster_grid = 0;  // workgroups of the persistent rasterizer
};

void drop_graph(BatchState* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->captured = false;
}

void destroy_batch(gs3d_ctx* c) {
    BatchState* b = c->batch;
    if (!b) return;
    for (int k = 0; k < kBatchBranches; k++) {
        if (b->streams[k]) (void)hipStreamSynchronize(b->streams[k]);
    }
    if (b->raster_stream) (void)hipStreamSynchronize(b->raster_stream);
    drop_graph(b);
    for (auto& g : b->group_done)
        for (auto& e : g)
            if (e) (void)hipEventDestroy(e);
    if (b->raster_join) (void)hipEventDestroy(b->raster_join);
    if (b->raster_stream) (void)hipStreamDestroy(b->raster_stream);
    for (int k = 0; k < kBatchBranches; k++) {
        if (b->join[k]) (void)hipEventDestroy(b->join[k]);
        if (b->streams[k]) (void)hipStreamDestroy(b->streams[k]);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->pinned_k) (void)hipHostFree(b->pinned_k);
    delete b;
    c->batch = nullptr;
}

int batch_init(gs3d_ctx* c) {
    if (c->batch) return GS3D_OK;
    BatchState* b = new BatchState();
    c->batch = b;
    for (int k = 0; k < kBatchBranches; k++) {
        HIP_TRY(c, hipStreamCreateWithFlags(&b->streams[k], hipStreamNonBlocking));
        HIP_TRY(c, hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming));
    }
    HIP_TRY(c, hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    HIP_TRY(c, hipStreamCreateWithFlags(&b->raster_stream, hipStreamNonBlocking));
    HIP_TRY(c, hipEventCreateWithFlags(&b->raster_join, hipEventDisableTiming));
    for (auto& g : b->group_done)
        for (auto& e : g) HIP_TRY(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return GS3D_OK;
}

std::string branch_name(int k, const char* what) { return "batch_b" + std::to_string(k) + "_" + what; }

// Passes of one view: shifts of the depth digits, then the tile digits below
// bit_length(T - 1) (zero digits of every key are skipped, as in a frame).
uint32_t batch_plan(uint32_t T, uint32_t bits, uint32_t* shifts) {
    const uint32_t tile_bits = T > 1 ? bit_length(T - 1) : 0;
    uint32_t p = 0;
    for (uint32_t sft = 0; sft < 32u + tile_bits; sft += bits) shifts[p++] = sft;
    return p;
}

// Allocate every batch buffer for the current views and capacities. Returns
// true in *moved when any block was (re)allocated (a captured graph then
// holds stale addresses).
int batch_buffers(gs3d_ctx* c, bool* moved) {
    BatchState* b = c->batch;
    const uint32_t n = c->n;
    const size_t V = b->views.size();
    size_t img = 0, to = 0;
    for (const auto& v : b->views) {
        img = std::max(img, v.img_off + (size_t)v.W * v.H * 4);
        to = std::max(to, v.to_off + (size_t)v.tiles_x * v.tiles_y + 1);
    }
    const uint32_t radix = 1u << c->digit_bits;
    // each view's sorted splat ids persist until the batch's one raster launch;
    // a view's slice holds its branch's key capacity
    b->val_total = 0;
    b->n_items = 0;
    b->max_tiles = 0;
    for (auto& v : b->views) {
        v.val_off = b->val_total;
        b->val_total += b->cap[v.branch];
        b->n_items += v.tiles_x * v.tiles_y;
        b->max_tiles = std::max(b->max_tiles, v.tiles_x * v.tiles_y);
    }
    std::vector<std::pair<std::string, size_t>> want = {
        {"batch_views", V * sizeof(ViewParams)},
        {"batch_sorted_values", b->val_total * 4},
        {"batch_raster_views", V * sizeof(RasterView)},
        {"batch_raster_queue", (size_t)b->n_items * sizeof(uint2)},
        {"batch_raster_ctl", (size_t)kMaxRasterGroups * kRasterQueueWords * 4},
        {"batch_projected", V * (size_t)n * kBatchRecordBytes},
        {"batch_depth_keys", V * batch_view_stride(n) * 4},
        {"batch_rects", V * batch_view_stride(n) * 8},
        {"batch_counters", V * kCtrTotalWords * 4},
        {"batch_tile_offsets", to * 4},
        {"batch_images", img * 2},
        {"batch_epoch", 4},
    };
    if (b->learned) {
        want.push_back({"batch_view_colors", V * (size_t)n * 8});
        want.push_back({"batch_eyes", V * 16});
    }
    for (int k = 0; k < kBatchBranches; k++) {
        const uint64_t cap = b->cap[k];
        want.push_back({branch_name(k, "depth_keys_b"), (size_t)n * 4});
        want.push_back({branch_name(k, "order_a"), (size_t)n * 4});
        want.push_back({branch_name(k, "order_b"), (size_t)n * 4});
        want.push_back({branch_name(k, "records"), ((size_t)n + 1) * 16});
        want.push_back({branch_name(k, "hist_partial"),
                        (size_t)std::max(emit_grid(n), key_hist_rows(n)) * kMaxPasses * radix * 4});
        want.push_back({branch_name(k, "keys_a"), cap * 8});
        want.push_back({branch_name(k, "keys_b"), cap * 8});
        want.push_back({branch_name(k, "values_a"), cap * 4});
        want.push_back({branch_name(k, "values_b"), cap * 4});
    }
    *moved = false;
    for (auto& w : want) {
        const void* before = c->bufs.count(w.first) ? c->bufs[w.first].ptr : nullptr;
        void* p;
        GS3D_TRY(ensure(c, w.first.c_str(), w.second, &p));
        if (p != before) {
            *moved = true;
            if (w.first == "batch_epoch") HIP_TRY(c, launch_fill_u32((uint32_t*)p, 0u, 1, c->stream));
        }
    }
    for (int k = 0; k < kBatchBranches; k++) {
        uint64_t* st;
        const void* before[2] = {c->bufs.count(branch_name(k, "scan_status")) ? c->bufs[branch_name(k, "scan_status")].ptr : nullptr,
                                 c->bufs.count(branch_name(k, "radix_status")) ? c->bufs[branch_name(k, "radix_status")].ptr : nullptr};
        GS3D_TRY(ensure_status(c, branch_name(k, "scan_status").c_str(), (size_t)n / kScanTile + 2, &st));
        if ((void*)st != before[0]) *moved = true;
        GS3D_TRY(ensure_status(c, branch_name(k, "radix_status").c_str(),
                               (size_t)radix_status_tiles((uint32_t)std::max<uint64_t>(b->cap[k], n)) * radix, &st));
        if ((void*)st != before[1]) *moved = true;
    }
    if (b->pinned_words < V) {
        if (b->pinned_k) HIP_TRY(c, hipHostFree(b->pinned_k));
        b->pinned_k = nullptr;
        HIP_TRY(c, hipHostMalloc((void**)&b->pinned_k, V * 4, hipHostMallocMapped));
        b->pinned_words = (uint32_t)V;
        *moved = true;
    }
    return GS3D_OK;
}

// Zero the replay base and every branch's lookback status words (on the
// context's stream, outside any capture; the branches are idle between calls).
int batch_reset_epochs(gs3d_ctx* c) {
    BatchState* b = c->batch;
    HIP_TRY(c, launch_fill_u32(buf<uint32_t>(c, "batch_epoch"), 0u, 1, c->stream));
    for (int k = 0; k < kBatchBranches; k++) {
        for (const char* what : {"scan_status", "radix_status"}) {
            const DevBuf& d = c->bufs[branch_name(k, what)];
            if (d.ptr) HIP_TRY(c, launch_fill_u32(static_cast<uint32_t*>(d.ptr), 0u, d.cap / 4, c->stream));
        }
    }
    b->epoch_bumps = 0;
    return GS3D_OK;
}

// LSD passes of one view on its branch (the frame's run_passes with the
// batch's status block and replay epochs).
int batch_passes(gs3d_ctx* c, hipStream_t s, void* ka, uint32_t* va, void* kb, uint32_t* vb,
                 uint32_t key_bytes, bool first_iota, const uint32_t* n_dev, uint64_t n_max,
                 const uint32_t* shifts, uint32_t passes, const uint32_t* hist, uint32_t* tickets,
                 uint64_t* status, const uint32_t* epoch_base, uint32_t* epoch_ctr, bool* in_b,
                 uint32_t key_bits, uint32_t key_tile, uint32_t* carry_a = nullptr, uint32_t* carry_b = nullptr,
                 uint32_t* last_values_out = nullptr) {
    // last_values_out (or null): where the last pass writes its values instead
    // of the ping-pong buffer
    const uint32_t radix = 1u << c->digit_bits;
    for (uint32_t p = 0; p < passes; p++) {
        const bool even = (p % 2) == 0;
        uint32_t* vout = (p + 1 == passes && last_values_out) ? last_values_out : (even ? vb : va);
        HIP_TRY(c, launch_radix_pass(even ? ka : kb, even ? va : vb, even ? kb : ka, vout, key_bytes,
                                     first_iota && p == 0, n_dev, (uint32_t)n_max, hist + p * radix, status,
                                     tickets + p, shifts[p], c->digit_bits, ++*epoch_ctr,
                                     carry_a ? (even ? carry_a : carry_b) : nullptr,
                                     carry_a ? (even ? carry_b : carry_a) : nullptr, s, epoch_base, key_tile,
                                     pass_match_bits(shifts[p], c->digit_bits, key_bits), carry_a != nullptr));
    }
    *in_b = (passes % 2) == 1;
    return GS3D_OK;
}

// The persistent rasterizer of views [v0, v1) on raster_stream: their tiles
// queued longest list first, then one launch that drains the queue (a fill
// kernel clears the group's counters: no memset node).
int batch_raster_group(gs3d_ctx* c, uint32_t group, uint32_t v0, uint32_t v1) {
    BatchState* b = c->batch;
    size_t item0 = 0;
    for (uint32_t i = 0; i < v0; i++) item0 += (size_t)b->views[i].tiles_x * b->views[i].tiles_y;
    uint32_t items = 0, max_tiles = 0;
    for (uint32_t i = v0; i < v1; i++) {
        const uint32_t t = b->views[i].tiles_x * b->views[i].tiles_y;
        items += t;
        max_tiles = std::max(max_tiles, t);
    }
    const RasterView* rviews = buf<RasterView>(c, "batch_raster_views") + v0;
    uint2* queue = buf<uint2>(c, "batch_raster_queue") + item0;
    uint32_t* qctl = buf<uint32_t>(c, "batch_raster_ctl") + (size_t)group * kRasterQueueWords;
    hipStream_t s = b->raster_stream;
    HIP_TRY(c, launch_fill_u32(qctl, 0u, kRasterQueueWords, s));
    HIP_TRY(c, launch_raster_queue_build(rviews, v1 - v0, max_tiles, qctl, queue, s));
    HIP_TRY(c, launch_rasterize_queue(rviews, queue, items, qctl, c->tile_x, c->raster_ppl, b->raster_grid,
                                      b->wave_items != 0, s));
    return GS3D_OK;
}

// Every launch of one batch, on c->stream and the branch streams.
int batch_enqueue(gs3d_ctx* c) {
    BatchState* b = c->batch;
    const uint32_t n = c->n;
    const uint32_t V = (uint32_t)b->views.size();
    const uint32_t bits = c->digit_bits, radix = 1u << bits;
    ViewParams* vparams = buf<ViewParams>(c, "batch_views");
    auto* projected = buf<float4>(c, "batch_projected");
    uint32_t* dk = buf<uint32_t>(c, "batch_depth_keys");
    uint2* rects = buf<uint2>(c, "batch_rects");
    uint32_t* counters = buf<uint32_t>(c, "batch_counters");
    uint32_t* tile_offsets = buf<uint32_t>(c, "batch_tile_offsets");
    uint16_t* images = buf<uint16_t>(c, "batch_images");
    uint32_t* epoch_base = buf<uint32_t>(c, "batch_epoch");
    uint32_t* sorted_values = buf<uint32_t>(c, "batch_sorted_values");

    HIP_TRY(c, launch_bump_epoch(epoch_base, c->stream));
    MultiPreprocessArgs ma{};
    ma.views = vparams;
    ma.n_views = V;
    if (b->learned) {
        LearnedColorArgs la{};
        la.params = buf<float>(c, "learned_params");
        la.latents = buf<uint16_t>(c, "learned_latents_f16");
        la.pos = buf<float4>(c, "splat_positions");
        la.eyes = buf<float>(c, "batch_eyes");
        la.colors = buf<uint2>(c, "batch_view_colors");
        la.n = n;
        la.n_views = V;
        HIP_TRY(c, launch_learned_colors(la, c->stream));
        ma.view_colors = la.colors;
    }
    ma.pos = buf<float4>(c, "splat_positions");
    ma.scale = buf<float4>(c, "splat_scales");
    ma.rot = buf<float4>(c, "splat_rotations");
    ma.color = buf<float4>(c, "splat_colors");
    ma.sh = c->sh_degree > 0 ? buf<float>(c, "splat_sh") : nullptr;
    ma.n = n;
    ma.sh_degree = c->sh_degree;
    ma.counter_words = (uint32_t)kCtrTotalWords;
    ma.count_slot = kCtrSplatCount;
    HIP_TRY(c, launch_preprocess_views(ma, c->stream));
    HIP_TRY(c, hipEventRecord(b->fork, c->stream));
    for (int k = 0; k < kBatchBranches; k++) HIP_TRY(c, hipStreamWaitEvent(b->streams[k], b->fork, 0));

    // view groups: contiguous view ranges of about equal pixel count, each
    // rasterized by one persistent launch on raster_stream once its views'
    // chains are done, while later views' chains still run
    uint32_t group_begin[kMaxRasterGroups], group_end[kMaxRasterGroups];
    {
        uint64_t total_px = 0;
        for (const auto& v : b->views) total_px += (uint64_t)v.W * v.H;
        uint32_t g = 0;
        uint64_t run = 0;
        group_begin[0] = 0;
        for (uint32_t vi = 0; vi < V; vi++) {
            run += (uint64_t)b->views[vi].W * b->views[vi].H;
            if (g + 1 < b->groups && run * b->groups >= total_px * (g + 1) && vi + 1 < V) {  // groups >= 2
                group_end[g] = vi + 1;
                group_begin[++g] = vi + 1;
            }
        }
        group_end[g] = V;
        for (uint32_t k = g + 1; k < kMaxRasterGroups; k++) group_begin[k] = group_end[k] = V + 1;
    }
    uint32_t group = 0;
    HIP_TRY(c, hipStreamWaitEvent(b->raster_stream, b->fork, 0));
    uint32_t epoch_ctr[kBatchBranches] = {};
    for (uint32_t vi = 0; vi < V; vi++) {
        const BatchView& v = b->views[vi];
        const int k = v.branch;
        hipStream_t s = b->streams[k];
        const uint64_t cap = b->cap[k];
        const uint32_t T = v.tiles_x * v.tiles_y;
        uint32_t shifts[kMaxPasses];
        const uint32_t passes = batch_plan(T, bits, shifts);
        const uint32_t depth_passes = 32u / bits, tile_passes = passes - depth_passes;
        uint32_t* ctr = counters + (size_t)vi * kCtrTotalWords;
        uint32_t* hist_depth = ctr + kCtrHist;
        uint32_t* hist_tile = ctr + kCtrHist + depth_passes * radix;
        uint32_t* hist_partial = buf<uint32_t>(c, branch_name(k, "hist_partial").c_str());
        const size_t ns = batch_view_stride(n);
        uint32_t* dka = dk + (size_t)vi * ns;
        uint32_t* dkb = buf<uint32_t>(c, branch_name(k, "depth_keys_b").c_str());
        uint32_t* ida = buf<uint32_t>(c, branch_name(k, "order_a").c_str());
        uint32_t* idb = buf<uint32_t>(c, branch_name(k, "order_b").c_str());
        uint4* recs = buf<uint4>(c, branch_name(k, "records").c_str());
        uint64_t* ka = buf<uint64_t>(c, branch_name(k, "keys_a").c_str());
        uint64_t* kb = buf<uint64_t>(c, branch_name(k, "keys_b").c_str());
        uint32_t* va = buf<uint32_t>(c, branch_name(k, "values_a").c_str());
        uint32_t* vb = buf<uint32_t>(c, branch_name(k, "values_b").c_str());
        uint64_t* scan_status = buf<uint64_t>(c, branch_name(k, "scan_status").c_str());
        uint64_t* radix_status = buf<uint64_t>(c, branch_name(k, "radix_status").c_str());

        HIP_TRY(c, launch_key_histogram(dka, 4, ctr + kCtrSplatCount, n, hist_depth, hist_partial, bits,
                                        depth_passes, 0, s));
        bool order_in_b = false;
        // packed rectangles (as set up for k_preprocess_views) move with the depth keys
        const bool packed = v.tiles_x <= 256u && v.tiles_y <= 256u;
        uint32_t* r8a = packed ? reinterpret_cast<uint32_t*>(rects + (size_t)vi * ns) : nullptr;
        uint32_t* r8b = packed ? r8a + ns : nullptr;
        GS3D_TRY(batch_passes(c, s, dka, ida, dkb, idb, 4, true, ctr + kCtrSplatCount, n, shifts, depth_passes,
                              hist_depth, ctr + kCtrSortTicket, radix_status, epoch_base, &epoch_ctr[k],
                              &order_in_b, 32, c->sort_key_tile32, r8a, r8b));
        ScanRectsArgs sa{};
        sa.order = order_in_b ? idb : ida;
        sa.depth_keys = order_in_b ? dkb : dka;
        sa.rects = packed ? nullptr : rects + (size_t)vi * ns;
        sa.rects8 = packed ? (order_in_b ? r8b : r8a) : nullptr;
        sa.recs = recs;
        sa.ref_counter = ctr + kCtrRefCounter;
        sa.status = scan_status;
        sa.ticket = ctr + kCtrScanTicket;
        sa.n = n;
        sa.tiles_x = v.tiles_x;
        sa.epoch = ++epoch_ctr[k];
        sa.epoch_base = epoch_base;
        HIP_TRY(c, launch_scan_rects(sa, s));
        EmitArgs ea{};
        ea.projected = nullptr;  // emit reads the scan's records
        ea.order = sa.order;
        ea.recs = recs;
        const uint32_t tb = tile_id_bytes(T);
        ea.tiles = tb == 4 ? reinterpret_cast<uint32_t*>(ka) : nullptr;
        ea.tiles16 = tb == 2 ? reinterpret_cast<uint16_t*>(ka) : nullptr;
        // with no tile-id pass (one tile) the emit order is the sorted order
        uint32_t* view_values = sorted_values + v.val_off;
        ea.values = tile_passes ? va : view_values;
        ea.hist = tile_passes ? hist_tile : nullptr;
        ea.hist_partial = tile_passes ? hist_partial : nullptr;
        ea.n = n;
        ea.tiles_x = v.tiles_x;
        ea.capacity = cap;
        ea.n_passes = tile_passes;
        ea.digit_bits = bits;
        ea.hist_shift0 = 32;
        HIP_TRY(c, launch_emit(ea, s));
        bool in_b = false;
        uint32_t tile_shifts[kMaxPasses];
        for (uint32_t p = 0; p < tile_passes; p++) tile_shifts[p] = shifts[depth_passes + p] - 32u;
        GS3D_TRY(batch_passes(c, s, ka, va, kb, vb, tb, false, ctr + kCtrRefCounter, cap, tile_shifts,
                              tile_passes, hist_tile, ctr + kCtrSortTicket + depth_passes, radix_status,
                              epoch_base, &epoch_ctr[k], &in_b, T > 1 ? bit_length(T - 1) : 0u, tile_key_tile(c),
                              nullptr, nullptr, view_values));
        uint32_t* to = tile_offsets + v.to_off;
        HIP_TRY(c, launch_identify_tiles(in_b ? kb : ka, tb, ctr + kCtrRefCounter, (uint32_t)cap, to, T,
                                         b->pinned_k + vi, s));
        if (b->groups == 0) {
            // per-view raster launches: each view's raster overlaps other views' chains
            // (mode 2 takes quarter lists: a batch's views are captured before their K is known)
            HIP_TRY(c, launch_rasterize_records(projected + (size_t)vi * n * 3, 3, view_values, to, c->tile_x,
                                                c->tile_y, v.tiles_x, v.tiles_y, v.W, v.H, c->raster_ppl,
                                                images + v.img_off, s, c->raster_quarters != 0));
        } else if (vi + 1 == group_end[group]) {
            // views [group_begin, vi] are queued on their branches: the group's
            // persistent raster waits for the branches up to here
            for (int kk = 0; kk < kBatchBranches; kk++) {
                HIP_TRY(c, hipEventRecord(b->group_done[group][kk], b->streams[kk]));
                HIP_TRY(c, hipStreamWaitEvent(b->raster_stream, b->group_done[group][kk], 0));
            }
            GS3D_TRY(batch_raster_group(c, group, group_begin[group], vi + 1));
            group++;
        }
    }
    for (int k = 0; k < kBatchBranches; k++) {
        if (epoch_ctr[k] >= kEpochStride)  // rebin_batch checks this before anything is queued
            return fail(c, GS3D_E_STATE, "replay epochs of branch %d exceed the stride", k);
        HIP_TRY(c, hipEventRecord(b->join[k], b->streams[k]));
        HIP_TRY(c, hipStreamWaitEvent(c->stream, b->join[k], 0));
    }
    HIP_TRY(c, hipEventRecord(b->raster_join, b->raster_stream));
    HIP_TRY(c, hipStreamWaitEvent(c->stream, b->raster_join, 0));
    return GS3D_OK;
}

int batch_capture(gs3d_ctx* c) {
    BatchState* b = c->batch;
    drop_graph(b);
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_enqueue(c);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != GS3D_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    HIP_TRY(c, e);
    b->graph = g;
    HIP_TRY(c, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->captured = true;
    return GS3D_OK;
}

int rebin_batch(gs3d_ctx* c, const gs3d_view* views, uint32_t n_views, uint32_t flags) {
    GS3D_TRY(batch_init(c));
    BatchState* b = c->batch;
    // layout: images and tile ranges back to back; branches by pixel count, longest first
    b->views.assign(n_views, BatchView{});
    std::vector<uint64_t> shape;
    size_t img = 0, to = 0;
    for (uint32_t i = 0; i < n_views; i++) {
        BatchView& v = b->views[i];
        v.cam = views[i].camera;
        v.W = views[i].width;
        v.H = views[i].height;
        v.tiles_x = (v.W + c->tile_x - 1) / c->tile_x;
        v.tiles_y = (v.H + c->tile_y - 1) / c->tile_y;
        v.img_off = img;
        v.to_off = to;
        img += (size_t)v.W * v.H * 4;
        img = (img + 63) & ~size_t(63);  // 128-B aligned image starts
        to += (size_t)v.tiles_x * v.tiles_y + 1;
        shape.push_back(((uint64_t)v.W << 32) | v.H);
    }
    std::vector<uint32_t> order(n_views);
    for (uint32_t i = 0; i < n_views; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        return (uint64_t)b->views[x].W * b->views[x].H > (uint64_t)b->views[y].W * b->views[y].H;
    });
    uint64_t load[kBatchBranches] = {};
    for (uint32_t i : order) {
        int best = 0;
        for (int k = 1; k < kBatchBranches; k++)
            if (load[k] < load[best]) best = k;
        b->views[i].branch = best;
        load[best] += (uint64_t)b->views[i].W * b->views[i].H;
    }
    // replay epochs each branch's launches take (a scan and the passes per view);
    // checked before anything is queued
    uint64_t epochs_needed[kBatchBranches] = {};
    for (const BatchView& v : b->views) {
        uint32_t shifts[kMaxPasses];
        epochs_needed[v.branch] += 1u + batch_plan(v.tiles_x * v.tiles_y, c->digit_bits, shifts);
    }
    for (int k = 0; k < kBatchBranches; k++)
        if (epochs_needed[k] >= kEpochStride)
            return fail(c, GS3D_E_INVALID_ARG, "batch needs %llu replay epochs on one stream (at most %u)",
                        (unsigned long long)epochs_needed[k], kEpochStride - 1u);
    const bool learned = (flags & GS3D_BATCH_LEARNED) != 0 && c->learned_init && c->learned_n == c->n &&
                         (c->learned_enabled || c->learned_force);
    const bool same = b->captured && shape == b->shape && b->n == c->n && b->deg == c->sh_degree &&
                      b->tile == c->tile_x && b->bits == c->digit_bits && b->ppl == c->raster_ppl &&
                      b->groups == std::min<uint32_t>(c->batch_raster_groups, n_views) &&
                      b->wg_per_cu == c->batch_raster_wg_per_cu && b->wave_items == c->batch_raster_wave_items &&
                      b->learned == learned;
    if (!same) {
        drop_graph(b);
        for (int k = 0; k < kBatchBranches; k++)
            if (b->cap[k] == 0) b->cap[k] = key_capacity_for(std::max<uint64_t>(8ull * c->n, 1u << 16));
    }
    b->shape = shape;
    b->n = c->n;
    b->deg = c->sh_degree;
    b->tile = c->tile_x;
    b->bits = c->digit_bits;
    b->ppl = c->raster_ppl;
    b->groups = std::min<uint32_t>(c->batch_raster_groups, n_views);
    b->wg_per_cu = c->batch_raster_wg_per_cu;
    b->wave_items = c->batch_raster_wave_items;
    b->learned = learned;
    b->have = false;
    b->splat_gen = c->splat_gen;
    const bool eager = (flags & GS3D_BATCH_EAGER) != 0;
    HIP_TRY(c, rasterize_queue_grid(c->tile_x, c->raster_ppl, c->batch_raster_wg_per_cu,
                                    c->batch_raster_wave_items != 0, &b->raster_grid));
    for (int attempt = 0; attempt < 3; attempt++) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        bool moved = false;
        GS3D_TRY(batch_buffers(c, &moved));
        if (moved) drop_graph(b);
        // cameras and per-view buffer addresses
        std::vector<ViewParams> vp(n_views);
        auto* projected = buf<float4>(c, "batch_projected");
        uint32_t* dk = buf<uint32_t>(c, "batch_depth_keys");
        uint2* rects = buf<uint2>(c, "batch_rects");
        uint32_t* counters = buf<uint32_t>(c, "batch_counters");
        for (uint32_t i = 0; i < n_views; i++) {
            vp[i].cam = views[i].camera;
            vp[i].tiles_x = b->views[i].tiles_x;
            vp[i].tiles_y = b->views[i].tiles_y;
            vp[i].inv_tile_x = (float)(1.0 / (double)c->tile_x);
            vp[i].inv_tile_y = (float)(1.0 / (double)c->tile_y);
            vp[i].out = projected + (size_t)i * c->n * 3;
            const size_t ns = batch_view_stride(c->n);
            vp[i].depth_keys = dk + (size_t)i * ns;
            // a view's 8 B per splat of rectangle memory holds, when packed, the two
            // 4-B arrays its depth passes carry the rectangles between
            const bool packed = b->views[i].tiles_x <= 256u && b->views[i].tiles_y <= 256u;
            vp[i].rects = packed ? nullptr : rects + (size_t)i * ns;
            vp[i].rects8 = packed ? reinterpret_cast<uint32_t*>(rects + (size_t)i * ns) : nullptr;
            vp[i].counters = counters + (size_t)i * kCtrTotalWords;
        }
        HIP_TRY(c, hipMemcpyAsync(buf<void>(c, "batch_views"), vp.data(), vp.size() * sizeof(ViewParams),
                                  hipMemcpyHostToDevice, c->stream));
        std::vector<RasterView> rv(n_views);
        uint32_t* sorted_values = buf<uint32_t>(c, "batch_sorted_values");
        uint32_t* tile_offsets = buf<uint32_t>(c, "batch_tile_offsets");
        uint16_t* images = buf<uint16_t>(c, "batch_images");
        for (uint32_t i = 0; i < n_views; i++) {
            const BatchView& v = b->views[i];
            rv[i].proj = projected + (size_t)i * c->n * 3;
            rv[i].values = sorted_values + v.val_off;
            rv[i].tile_offsets = tile_offsets + v.to_off;
            rv[i].image = reinterpret_cast<uint2*>(images + v.img_off);
            rv[i].tiles_x = v.tiles_x;
            rv[i].tiles = v.tiles_x * v.tiles_y;
            rv[i].width = v.W;
            rv[i].height = v.H;
        }
        HIP_TRY(c, hipMemcpyAsync(buf<void>(c, "batch_raster_views"), rv.data(), rv.size() * sizeof(RasterView),
                                  hipMemcpyHostToDevice, c->stream));
        std::vector<float> eyes;
        if (learned) {
            eyes.resize((size_t)n_views * 4);
            for (uint32_t i = 0; i < n_views; i++)
                for (int k = 0; k < 4; k++) eyes[(size_t)i * 4 + k] = views[i].camera.position[k];
            HIP_TRY(c, hipMemcpyAsync(buf<void>(c, "batch_eyes"), eyes.data(), eyes.size() * 4,
                                      hipMemcpyHostToDevice, c->stream));
        }
        if (b->epoch_bumps + 1u >= kEpochBumpsPerRound) {
            // the replay base would wrap: start a new round of bases from zeroed status words
            GS3D_TRY(batch_reset_epochs(c));
        }
        b->epoch_bumps++;
        if (eager) {
            GS3D_TRY(batch_enqueue(c));
        } else {
            if (!b->captured) GS3D_TRY(batch_capture(c));
            HIP_TRY(c, hipGraphLaunch(b->exec, c->stream));
            b->replays++;
        }
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        b->K.assign(n_views, 0);
        bool grow = false;
        uint64_t need[kBatchBranches] = {};
        for (uint32_t i = 0; i < n_views; i++) {
            b->K[i] = b->pinned_k[i];
            const int k = b->views[i].branch;
            need[k] = std::max(need[k], b->K[i]);
            if (b->K[i] > b->cap[k]) grow = true;
        }
        if (!grow) {
            b->have = true;
            return GS3D_OK;
        }
        for (int k = 0; k < kBatchBranches; k++)
            if (need[k] > b->cap[k]) b->cap[k] = key_capacity_for(need[k] + need[k] / 4 + kSortTile);
        drop_graph(b);
    }
    return fail(c, GS3D_E_STATE, "batch key capacity did not settle");
}

// ---- learned colour ------------------------------------------------------------

constexpr int kLearnedParams = GS3D_LEARNED_PARAMS;

// The 256 training directions: a Fibonacci sphere (z evenly spaced, azimuth by
// the golden angle), computed in f64 and stored as f32.
void learned_directions(float* dirs) {
    const double golden = 3.14159265358979323846 * (3.0 - std::sqrt(5.0));
    for (int k = 0; k < GS3D_LEARNED_DIRS; k++) {
        const double z = 1.0 - (2.0 * k + 1.0) / GS3D_LEARNED_DIRS;
        const double r = std::sqrt(std::max(0.0, 1.0 - z * z));
        const double phi = golden * k;
        dirs[3 * k] = (float)(r * std::cos(phi));
        dirs[3 * k + 1] = (float)(r * std::sin(phi));
        dirs[3 * k + 2] = (float)z;
    }
}

// Real SH basis of degree 3 at a unit direction (the constants of the 3DGS
// formulation, as in the preprocess kernel), in f64.
void sh_basis_f64(double x, double y, double z, double* b) {
    const double xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
    b[0] = 0.28209479177387814;
    b[1] = -0.4886025119029199 * y;
    b[2] = 0.4886025119029199 * z;
    b[3] = -0.4886025119029199 * x;
    b[4] = 1.0925484305920792 * xy;
    b[5] = -1.0925484305920792 * yz;
    b[6] = 0.31539156525252005 * (2.0 * zz - xx - yy);
    b[7] = -1.0925484305920792 * xz;
    b[8] = 0.5462742152960396 * (xx - yy);
    b[9] = -0.5900435899266435 * y * (3.0 * xx - yy);
    b[10] = 2.890611442640554 * xy * z;
    b[11] = -0.4570457994644658 * y * (4.0 * zz - xx - yy);
    b[12] = 0.3731763325901154 * z * (2.0 * zz - 3.0 * xx - 3.0 * yy);
    b[13] = -0.4570457994644658 * x * (4.0 * zz - xx - yy);
    b[14] = 1.445305721320277 * z * (xx - yy);
    b[15] = -0.5900435899266435 * x * (xx - 3.0 * yy);
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double normal01(uint64_t& s) {
    const double u1 = ((splitmix(s) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
    const double u2 = (splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
}

int learned_require(gs3d_ctx* c) {
    if (!c->learned_init) return fail(c, GS3D_E_STATE, "learned colour not initialised (gs3d_learned_init)");
    if (c->learned_n != c->n) return fail(c, GS3D_E_STATE, "splats changed since gs3d_learned_init");
    return GS3D_OK;
}

int learned_buffers(gs3d_ctx* c) {
    const size_t n8 = (size_t)c->n * 8;
    const uint32_t waves = learned_train_waves(c->n);
    const std::pair<const char*, size_t> want[] = {
        {"learned_params", kLearnedParams * 4}, {"learned_pm", kLearnedParams * 4}, {"learned_pv", kLearnedParams * 4},
        {"learned_grad", kLearnedParams * 4},   {"learned_partial", (size_t)waves * kLearnedParams * 4},
        {"learned_loss_partial", waves * 4u},   {"learned_loss", 16 * 4},
        {"learned_latents", n8 * 4},            {"learned_lm", n8 * 4},  {"learned_lv", n8 * 4},
        {"learned_lgrad", n8 * 4},              {"learned_latents_f16", n8 * 2},
        {"learned_dirs", GS3D_LEARNED_DIRS * 3 * 4}, {"learned_basis", GS3D_LEARNED_DIRS * 16 * 4},
        {"learned_reduce_scratch", learned_reduce_scratch_floats() * 4},
    };
    for (const auto& w : want) {
        void* p;
        GS3D_TRY(ensure(c, w.first, w.second, &p));
    }
    return GS3D_OK;
}

// One forward/backward over all splats x directions [first, first + 16):
// gradients into learned_grad / learned_lgrad, the sum of squared errors into
// learned_loss[slot].
int learned_step_grads(gs3d_ctx* c, uint32_t first_dir, uint32_t slot) {
    LearnedTrainArgs ta{};
    ta.params = buf<float>(c, "learned_params");
    ta.latents = buf<float>(c, "learned_latents");
    ta.sh = c->sh_degree > 0 ? buf<float>(c, "splat_sh") : nullptr;
    ta.sh_degree = c->sh_degree;
    ta.dirs = buf<float>(c, "learned_dirs");
    ta.basis = buf<float>(c, "learned_basis");
    ta.first_dir = first_dir;
    ta.n = c->n;
    ta.grad_scale = 1.0f / (3.0f * 16.0f * (float)std::max(1u, c->n));
    ta.grad_partial = buf<float>(c, "learned_partial");
    ta.grad_latents = buf<float>(c, "learned_lgrad");
    ta.loss_partial = buf<float>(c, "learned_loss_partial");
    const uint32_t waves = learned_train_waves(c->n);
    HIP_TRY(c, launch_learned_train(ta, waves, c->stream));
    HIP_TRY(c, launch_learned_reduce(ta.grad_partial, waves, buf<float>(c, "learned_grad"), ta.loss_partial,
                                     buf<float>(c, "learned_loss") + slot, buf<float>(c, "learned_reduce_scratch"),
                                     c->stream));
    return GS3D_OK;
}

// Used by ply.cpp, which has no view of the context's layout.
__attribute__((visibility("hidden"))) int gs3d_internal_fail(gs3d_ctx* c, int code, const char* msg) {
    return fail(c, code, "%s", msg);
}
__attribute__((visibility("hidden"))) int gs3d_internal_set_device(gs3d_ctx* c) { return set_device(c); }

// ---- allocation-free sort: static scratch --------------------------------------
//
// gs3d_sort_pairs_noalloc keeps everything beyond the caller's inputs and
// outputs in `scratch`, a file-scope device array each GPU holds from module
// load to process end. Host code never passes the array's name to a kernel
// (in host code it is a stand-in, not a device address): gs3d_create looks up
// its address on each device once, with hipGetSymbolAddress, and the entry
// point uses that per-device address. Layout, 256-B aligned sections:
//   control   64 u32: [0] = n (the pass kernels read the key count on the
//             device), [1 + p] = pass p's workgroup ticket
//   hist      [kMaxPasses][256] u32 digit histograms of the input
//   partial   the histogram kernel's per-workgroup rows (<= 1,024 rows of at
//             most 8 passes x 256 or 16 x 16 digits)
//   status    lookback words of one pass over the maximum key count (u64,
//             tagged with a process-wide epoch, never cleared)
//   keys      the ping-pong side of the passes that is not the caller's output:
//   values    GS3D_SORT_NOALLOC_MAX_PAIRS u64 keys and u32 values
namespace {

constexpr size_t na_align(size_t x) { return (x + 255u) & ~size_t(255); }
constexpr size_t kNaMaxPairs = GS3D_SORT_NOALLOC_MAX_PAIRS;
constexpr size_t kNaCtlWords = 64;
constexpr size_t kNaHistOff = na_align(kNaCtlWords * 4);
constexpr size_t kNaPartialOff = na_align(kNaHistOff + (size_t)kMaxPasses * 256 * 4);
constexpr size_t kNaPartialWords = 1024u * 2048u;
constexpr size_t kNaStatusWords = ((kNaMaxPairs + kSortTile - 1) / kSortTile + 1) * 256;
constexpr size_t kNaStatusOff = na_align(kNaPartialOff + kNaPartialWords * 4);
constexpr size_t kNaKeysOff = na_align(kNaStatusOff + kNaStatusWords * 8);
constexpr size_t kNaValuesOff = na_align(kNaKeysOff + kNaMaxPairs * 8);
constexpr size_t kNaBytes = na_align(kNaValuesOff + kNaMaxPairs * 4);

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mutex;
uint8_t* g_na_scratch[kNaMaxDevices] = {};  // device address of `scratch` on each device

}  // namespace

static __device__ __align__(256) uint8_t scratch[kNaBytes];

namespace {

// With `device` current: find this device's copy of `scratch` (once per device).
hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= kNaMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (g_na_scratch[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_na_scratch[device] = static_cast<uint8_t*>(ptr);
    return hipSuccess;
}

uint8_t* noalloc_scratch(int device) {
    if (device < 0 || device >= kNaMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_na_mutex);
    return g_na_scratch[device];
}

}  // namespace

extern "C" {

int gs3d_abi_version(void) { return GS3D_ABI_VERSION; }

const char* gs3d_last_error(const gs3d_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

uint32_t gs3d_reference_pass_stages(const char* label) {
    if (!label) return 0u;
    struct Row {
        const char* label;
        uint32_t stages;
    };
    const uint32_t sort = (1u << GS3D_STAGE_DEPTH_SORT) | (1u << GS3D_STAGE_SORT);
    static const Row kRows[] = {
        {"preprocess-splats", 1u << GS3D_STAGE_PREPROCESS},
        {"prefix-scan-local", 1u << GS3D_STAGE_SCAN},
        {"prefix-scan-blocks", 1u << GS3D_STAGE_SCAN},
        {"prefix-scan-add", 1u << GS3D_STAGE_SCAN},
        {"emit-tile-refs", 1u << GS3D_STAGE_EMIT},
        {"radix-histogram", sort},
        {"scan-local", sort},
        {"scan-add", sort},
        {"radix-scatter", sort},
        {"identify-tile-ranges", 1u << GS3D_STAGE_IDENTIFY},
        {"rasterize-tiles", 1u << GS3D_STAGE_RASTER},
        {"tile-overlay", 1u << GS3D_STAGE_COMPOSITE},
        {"composite-splats", 1u << GS3D_STAGE_COMPOSITE},
    };
    for (const Row& r : kRows)
        if (!std::strcmp(label, r.label)) return r.stages;
    return 0u;
}

const char* gs3d_stage_label(int stage) {
    return (stage >= 0 && stage < GS3D_STAGE__COUNT) ? kStageLabels[stage] : "";
}

int gs3d_create(int device, gs3d_ctx** out) {
    if (!out) return GS3D_E_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return GS3D_E_NO_DEVICE;
    if (device < 0 || device >= count) return GS3D_E_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return GS3D_E_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return GS3D_E_NO_DEVICE;
    gs3d_ctx* c = new gs3d_ctx();
    c->device = device;
    c->limits[0] = (uint32_t)prop.maxThreadsPerBlock;
    c->limits[1] = (uint32_t)prop.maxThreadsDim[0];
    c->limits[2] = (uint32_t)prop.maxThreadsDim[1];
    c->limits[3] = (uint32_t)prop.maxThreadsDim[2];
    c->limits[4] = (uint32_t)prop.maxGridSize[0];
    c->limits[5] = (uint32_t)prop.sharedMemPerBlock;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&c->pinned, 64, hipHostMallocMapped) != hipSuccess ||
        noalloc_scratch_init(device) != hipSuccess) {
        gs3d_destroy(c);
        return GS3D_E_HIP;
    }
    for (int r = 0; r < gs3d_ctx::kRing; r++)
        for (int s = 0; s < GS3D_STAGE__COUNT; s++)
            for (int w = 0; w < 2; w++)
                if (hipEventCreate(&c->ring[r][s][w]) != hipSuccess) {
                    gs3d_destroy(c);
                    return GS3D_E_HIP;
                }
    for (int w = 0; w < 2; w++)
        if (hipEventCreate(&c->composite_ev[w]) != hipSuccess) {
            gs3d_destroy(c);
            return GS3D_E_HIP;
        }
    *out = c;
    return GS3D_OK;
}

int gs3d_destroy(gs3d_ctx* c) {
    if (!c) return GS3D_E_INVALID_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->bufs)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    for (int r = 0; r < gs3d_ctx::kRing; r++)
        for (int s = 0; s < GS3D_STAGE__COUNT; s++)
            for (int w = 0; w < 2; w++)
                if (c->ring[r][s][w]) (void)hipEventDestroy(c->ring[r][s][w]);
    for (int w = 0; w < 2; w++)
        if (c->composite_ev[w]) (void)hipEventDestroy(c->composite_ev[w]);
    destroy_batch(c);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GS3D_OK;
}

int gs3d_sync(gs3d_ctx* c) {
    if (!c) return GS3D_E_INVALID_ARG;
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GS3D_OK;
}

int gs3d_device_limits(gs3d_ctx* c, uint32_t out[6]) {
    if (!c || !out) return GS3D_E_INVALID_ARG;
    std::memcpy(out, c->limits, sizeof c->limits);
    return GS3D_OK;
}

int gs3d_set_splats(gs3d_ctx* c, const float* pos4, const float* scale4, const float* rot4,
                    const float* color4, const float* f_dc, const float* sh_rest, int sh_degree,
                    uint32_t n) {
    if (!c) return GS3D_E_INVALID_ARG;
    if (sh_degree < 0 || sh_degree > 3) return fail(c, GS3D_E_INVALID_ARG, "sh_degree %d not in 0..3", sh_degree);
    if (n > 0 && (!pos4 || !scale4 || !rot4 || !color4))
        return fail(c, GS3D_E_INVALID_ARG, "null splat array");
    if (n > 0 && sh_degree > 0 && (!f_dc || !sh_rest))
        return fail(c, GS3D_E_INVALID_ARG, "sh_degree %d needs f_dc and sh_rest", sh_degree);
    if ((uint64_t)n * 16 > (1ull << 40)) return fail(c, GS3D_E_INVALID_ARG, "too many splats");
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    const char* names[4] = {"splat_positions", "splat_scales", "splat_rotations", "splat_colors"};
    const float* src[4] = {pos4, scale4, rot4, color4};
    for (int k = 0; k < 4; k++) {
        float4* d;
        GS3D_TRY(ensure_t(c, names[k], n, &d));
        if (n) HIP_TRY(c, hipMemcpyAsync(d, src[k], (size_t)n * 16, hipMemcpyHostToDevice, c->stream));
    }
    if (sh_degree > 0 && n > 0) {
        const int ncoef = (sh_degree + 1) * (sh_degree + 1);
        const size_t rest = (size_t)3 * (ncoef - 1);
        float *dc_dev, *rest_dev, *sh;
        GS3D_TRY(ensure_t(c, "staging_f_dc", (size_t)n * 3, &dc_dev));
        GS3D_TRY(ensure_t(c, "staging_sh_rest", (size_t)n * rest, &rest_dev));
        GS3D_TRY(ensure_t(c, "splat_sh", (size_t)n * 3 * ncoef, &sh));
        HIP_TRY(c, hipMemcpyAsync(dc_dev, f_dc, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(rest_dev, sh_rest, (size_t)n * rest * 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, launch_sh_relayout(dc_dev, rest_dev, sh, n, sh_degree, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->n = n;
    c->sh_degree = sh_degree;
    c->have_frame = false;
    c->splat_gen++;
    c->key_capacity = 0;
    // latents belong to the splats they were trained on
    c->learned_init = false;
    c->learned_enabled = false;
    return GS3D_OK;
}

uint32_t gs3d_splat_count(const gs3d_ctx* c) { return c ? c->n : 0; }

int gs3d_set_tile_size(gs3d_ctx* c, uint32_t x, uint32_t y, uint32_t out_xy[2]) {
    if (!c) return GS3D_E_INVALID_ARG;
    // clampTileSize (GaussianSplatRenderer.ts:162-167) against the device limits
    uint32_t tx = std::max<uint32_t>(1, std::min({x ? x : 1u, c->limits[1], c->limits[0]}));
    uint32_t ty = std::max<uint32_t>(1, std::min({y ? y : 1u, c->limits[2], c->limits[0] / tx}));
    if (!((tx == 8 && ty == 8) || (tx == 16 && ty == 16) || (tx == 32 && ty == 32)))
        return fail(c, GS3D_E_UNKNOWN_NAME, "no rasterize-tiles variant for a %ux%u tile (8, 16, 32 square)",
                    tx, ty);
    GS3D_TRY(settle(c));
    c->tile_x = tx;
    c->tile_y = ty;
    if (!raster_variant_exists(tx, c->raster_ppl)) c->raster_ppl = 1;  // 8x8 has one wavefront
    if (out_xy) {
        out_xy[0] = tx;
        out_xy[1] = ty;
    }
    return GS3D_OK;
}

int gs3d_set_kernel_variant(gs3d_ctx* c, const char* kernel, const char* param, int value) {
    if (!c || !kernel || !param) return GS3D_E_INVALID_ARG;
    GS3D_TRY(settle(c));
    if (!std::strcmp(kernel, "rasterize-tiles") && !std::strcmp(param, "PIXELS_PER_LANE")) {
        if (value < 0 || !raster_variant_exists(c->tile_x, (uint32_t)value))
            return fail(c, GS3D_E_UNKNOWN_NAME, "no rasterize-tiles variant with %d pixels per lane for %u px tiles",
                        value, c->tile_x);
        c->raster_ppl = (uint32_t)value;
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "rasterize-tiles") && !std::strcmp(param, "QUARTER_LISTS")) {
        if (value != 0 && value != 1 && value != 2)
            return fail(c, GS3D_E_UNKNOWN_NAME, "rasterize-tiles QUARTER_LISTS %d not 0, 1 or 2", value);
        if (c->raster_quarters != (uint32_t)value && c->batch) drop_graph(c->batch);
        c->raster_quarters = (uint32_t)value;
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "rasterize-tiles-batch") && !std::strcmp(param, "GROUPS")) {
        if (value < 0 || value > (int)kMaxRasterGroups)
            return fail(c, GS3D_E_UNKNOWN_NAME, "rasterize-tiles-batch GROUPS %d not in 0..%u", value,
                        kMaxRasterGroups);
        c->batch_raster_groups = (uint32_t)value;
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "rasterize-tiles-batch") && !std::strcmp(param, "WAVE_ITEMS")) {
        if (value != 0 && value != 1)
            return fail(c, GS3D_E_UNKNOWN_NAME, "rasterize-tiles-batch WAVE_ITEMS %d not 0 or 1", value);
        c->batch_raster_wave_items = (uint32_t)value;
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "rasterize-tiles-batch") && !std::strcmp(param, "WG_PER_CU")) {
        if (value < 0 || value > 16)
            return fail(c, GS3D_E_UNKNOWN_NAME, "rasterize-tiles-batch WG_PER_CU %d not in 0..16", value);
        c->batch_raster_wg_per_cu = (uint32_t)value;
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "radix-sort") && !std::strcmp(param, "RADIX_BITS")) {
        if (value != 4 && value != 8)
            return fail(c, GS3D_E_UNKNOWN_NAME, "no radix-sort variant with %d-bit digits", value);
        c->digit_bits = (uint32_t)value;
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "radix-sort") && !std::strcmp(param, "KEY_TILE")) {
        if (value != 0 && value != 8192 && value != 16384)
            return fail(c, GS3D_E_UNKNOWN_NAME, "no radix-sort variant with %d-key tiles (0, 8192 or 16384)", value);
        c->sort_key_tile = (uint32_t)value;
        if (c->batch) drop_graph(c->batch);
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "radix-sort-splats") && !std::strcmp(param, "KEY_TILE")) {
        if (value != 0 && value != 4096 && value != 8192)
            return fail(c, GS3D_E_UNKNOWN_NAME, "no radix-sort-splats variant with %d-key tiles (0, 4096 or 8192)", value);
        c->sort_key_tile32 = (uint32_t)value;
        if (c->batch) drop_graph(c->batch);
        return GS3D_OK;
    }
    return fail(c, GS3D_E_UNKNOWN_NAME, "no kernel variant %s.%s", kernel, param);
}

int gs3d_get_kernel_variant(gs3d_ctx* c, const char* kernel, const char* param, int* value) {
    if (!c || !kernel || !param || !value) return GS3D_E_INVALID_ARG;
    if (!std::strcmp(kernel, "rasterize-tiles") && !std::strcmp(param, "PIXELS_PER_LANE")) {
        *value = (int)c->raster_ppl;
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "rasterize-tiles") && !std::strcmp(param, "TILE_X")) {
        *value = (int)c->tile_x;
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "rasterize-tiles") && !std::strcmp(param, "QUARTER_LISTS")) {
        *value = (int)c->raster_quarters;
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "rasterize-tiles-batch") && !std::strcmp(param, "GROUPS")) {
        *value = (int)c->batch_raster_groups;
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "rasterize-tiles-batch") && !std::strcmp(param, "WG_PER_CU")) {
        *value = (int)c->batch_raster_wg_per_cu;
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "rasterize-tiles-batch") && !std::strcmp(param, "WAVE_ITEMS")) {
        *value = (int)c->batch_raster_wave_items;
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "radix-sort") && !std::strcmp(param, "RADIX_BITS")) {
        *value = (int)c->digit_bits;
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "radix-sort") && !std::strcmp(param, "KEY_TILE")) {
        *value = (int)c->sort_key_tile;
        return GS3D_OK;
    }
    if (!std::strcmp(kernel, "radix-sort-splats") && !std::strcmp(param, "KEY_TILE")) {
        *value = (int)c->sort_key_tile32;
        return GS3D_OK;
    }
    return fail(c, GS3D_E_UNKNOWN_NAME, "no kernel variant %s.%s", kernel, param);
}

int gs3d_get_tile_size(const gs3d_ctx* c, uint32_t out_xy[2]) {
    if (!c || !out_xy) return GS3D_E_INVALID_ARG;
    out_xy[0] = c->tile_x;
    out_xy[1] = c->tile_y;
    return GS3D_OK;
}

int gs3d_get_tile_count(const gs3d_ctx* c, uint32_t out_xy[2]) {
    if (!c || !out_xy) return GS3D_E_INVALID_ARG;
    out_xy[0] = c->tiles_x;
    out_xy[1] = c->tiles_y;
    return GS3D_OK;
}

int gs3d_rebin(gs3d_ctx* c, const gs3d_camera* camera, uint32_t width, uint32_t height,
               uint32_t flags) {
    if (!c || !camera) return GS3D_E_INVALID_ARG;
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return fail(c, GS3D_E_INVALID_ARG, "viewport %ux%u out of range", width, height);
    if (!c->bufs.count("splat_positions")) return fail(c, GS3D_E_STATE, "rebin before set_splats");
    GS3D_TRY(set_device(c));
    c->last_cam = *camera;
    c->last_w = width;
    c->last_h = height;
    c->last_flags = flags;
    c->have_frame = true;
    return enqueue_rebin(c, camera, width, height, flags);
}

int gs3d_ref_count(gs3d_ctx* c, uint64_t* out) {
    if (!c || !out) return GS3D_E_INVALID_ARG;
    if (!c->have_frame) return fail(c, GS3D_E_STATE, "no rebin yet");
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    *out = c->K;
    return GS3D_OK;
}

int gs3d_launch_geometry(const gs3d_ctx* c, const char* kernel, const uint64_t problem[3], uint32_t workgroup[3],
                         uint64_t grid[3]) {
    if (!kernel || !problem || !workgroup || !grid) return GS3D_E_INVALID_ARG;
    const uint32_t tile = c ? c->tile_x : 16u;
    const uint32_t ppl = c ? c->raster_ppl : 1u;
    const uint32_t key_tile32 = c ? c->sort_key_tile32 : 0u;
    const uint32_t key_tile = c ? tile_key_tile(c) : 16384u;
    const uint64_t n64 = problem[0];
    if (n64 > 0xFFFFFFFFull || problem[1] > 0xFFFFFFFFull) return GS3D_E_INVALID_ARG;
    const uint32_t n = (uint32_t)n64;
    const uint64_t tiles_x = (problem[0] + tile - 1) / tile, tiles_y = (problem[1] + tile - 1) / tile;
    LaunchShape sh{};
    uint64_t g[3] = {0, 1, 1};
    if (!std::strcmp(kernel, "preprocess-splats")) {
        sh = preprocess_launch_shape(n);
    } else if (!std::strcmp(kernel, "radix-histogram")) {
        sh = key_histogram_launch_shape(n);
    } else if (!std::strcmp(kernel, "radix-sort-splats")) {
        sh = radix_pass_launch_shape(4, key_tile32, n);
    } else if (!std::strcmp(kernel, "prefix-scan")) {
        sh = scan_rects_launch_shape(n);
    } else if (!std::strcmp(kernel, "emit-tile-refs")) {
        sh = emit_launch_shape(n);
    } else if (!std::strcmp(kernel, "radix-sort")) {
        sh = radix_pass_launch_shape(4, key_tile, n);
    } else if (!std::strcmp(kernel, "identify-tile-ranges")) {
        sh = identify_launch_shape((uint32_t)(tiles_x * tiles_y));
    } else if (!std::strcmp(kernel, "rasterize-tiles")) {
        sh = raster_launch_shape(tile, ppl, (uint32_t)(tiles_x * tiles_y));
    } else if (!std::strcmp(kernel, "composite-splats")) {
        sh = composite_launch_shape((uint32_t)(problem[0] * problem[1]));
    } else {
        return GS3D_E_UNKNOWN_NAME;
    }
    g[0] = sh.grid;
    workgroup[0] = sh.block;
    workgroup[1] = workgroup[2] = 1;
    for (int i = 0; i < 3; i++) grid[i] = g[i];
    return GS3D_OK;
}

int gs3d_sort_passes(gs3d_ctx* c, uint32_t* n_passes, uint32_t bit_offsets[16]) {
    if (!c || !n_passes) return GS3D_E_INVALID_ARG;
    *n_passes = c->n_passes;
    if (bit_offsets)
        for (uint32_t p = 0; p < c->n_passes; p++) bit_offsets[p] = c->shifts[p];
    return GS3D_OK;
}

int gs3d_read(gs3d_ctx* c, int id, void* dst, size_t bytes, size_t* written) {
    if (!c) return GS3D_E_INVALID_ARG;
    if (!c->have_frame) return fail(c, GS3D_E_STATE, "no rebin yet");
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    const uint64_t K = c->K;
    const size_t px = (size_t)c->width * c->height;
    const void* src = nullptr;
    size_t size = 0;
    bool widen = false;
    switch (id) {
        case GS3D_BUF_PROJECTED_SPLATS: src = buf<void>(c, "projected_splats"); size = (size_t)c->n * 64; break;
        case GS3D_BUF_SPLAT_REF_COUNTS: src = buf<void>(c, "splat_ref_counts"); size = (size_t)c->n * 4; break;
        case GS3D_BUF_SPLAT_REF_OFFSETS: {
            uint32_t* offs;
            GS3D_TRY(enqueue_id_order_scan(c, &offs));
            src = offs;
            size = ((size_t)c->n + 1) * 4;
            break;
        }
        case GS3D_BUF_REF_COUNTER: src = buf<uint32_t>(c, "counters") + kCtrRefCounter; size = 4; break;
        case GS3D_BUF_SORT_KEYS: {
            // the frame sorts tile ids; the (tile << 32 | depth) keys are rebuilt
            // from each key's splat
            uint64_t* rk;
            GS3D_TRY(ensure_t(c, "readback_keys", std::max<uint64_t>(K, 1), &rk));
            HIP_TRY(c, launch_rebuild_keys(buf<void>(c, c->result_in_b ? "sort_keys_b" : "sort_keys_a"),
                                           c->frame_tile_bytes,
                                           buf<uint32_t>(c, c->result_in_b ? "sort_values_b" : "sort_values_a"),
                                           buf<gs3d_projected_splat>(c, "projected_splats"),
                                           buf<uint32_t>(c, "counters") + kCtrRefCounter, (uint32_t)K, rk, c->stream));
            src = rk;
            size = K * 8;
            break;
        }
        case GS3D_BUF_SORT_VALUES: src = buf<void>(c, c->result_in_b ? "sort_values_b" : "sort_values_a"); size = K * 4; break;
        case GS3D_BUF_TILE_OFFSETS: src = buf<void>(c, "tile_offsets"); size = ((size_t)c->tiles_x * c->tiles_y + 1) * 4; break;
        case GS3D_BUF_IMAGE_RGBA16F: src = buf<void>(c, "splat_target"); size = px * 8; break;
        case GS3D_BUF_IMAGE_RGBA32F: src = buf<void>(c, "splat_target"); size = px * 16; widen = true; break;
        case GS3D_BUF_UNSORTED_KEYS:
            if (!c->keep_unsorted) return fail(c, GS3D_E_STATE, "rebin without GS3D_REBIN_KEEP_UNSORTED");
            src = buf<void>(c, "unsorted_keys"); size = K * 8; break;
        case GS3D_BUF_UNSORTED_VALUES:
            if (!c->keep_unsorted) return fail(c, GS3D_E_STATE, "rebin without GS3D_REBIN_KEEP_UNSORTED");
            src = buf<void>(c, "unsorted_values"); size = K * 4; break;
        case GS3D_BUF_DEBUG_TARGET_RGBA16F:
            if (c->debug_target_px != px) return fail(c, GS3D_E_STATE, "no composite in debug mode at this viewport");
            src = buf<void>(c, "debug_target"); size = px * 8; break;
        case GS3D_BUF_FRAMEBUFFER:
            if (c->framebuffer_px != px) return fail(c, GS3D_E_STATE, "no composite at this viewport");
            src = buf<void>(c, "framebuffer"); size = px * 4; break;
        default: return fail(c, GS3D_E_UNKNOWN_NAME, "unknown buffer id %d", id);
    }
    if (written) *written = size;
    if (!dst) return GS3D_OK;  // size query
    if (bytes < size) return fail(c, GS3D_E_TOO_SMALL, "buffer %d needs %zu bytes, got %zu", id, size, bytes);
    if (size == 0) return GS3D_OK;
    if (widen) {
        float* tmp;
        GS3D_TRY(ensure_t(c, "readback_f32", px * 4, &tmp));
        HIP_TRY(c, launch_f16_to_f32(static_cast<const uint16_t*>(src), tmp, px * 4, c->stream));
        src = tmp;
    }
    HIP_TRY(c, hipMemcpyAsync(dst, src, size, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GS3D_OK;
}

int gs3d_read_tile_counts(gs3d_ctx* c, uint32_t* counts, size_t n, size_t* written) {
    if (!c) return GS3D_E_INVALID_ARG;
    const size_t T = (size_t)c->tiles_x * c->tiles_y;
    if (written) *written = T;
    if (!c->have_frame) return fail(c, GS3D_E_STATE, "no rebin yet");
    if (!counts) return GS3D_OK;
    if (n < T) return fail(c, GS3D_E_TOO_SMALL, "need %zu counts, got %zu", T, n);
    std::vector<uint32_t> off(T + 1);
    GS3D_TRY(gs3d_read(c, GS3D_BUF_TILE_OFFSETS, off.data(), off.size() * 4, nullptr));
    for (size_t t = 0; t < T; t++) counts[t] = off[t + 1] >= off[t] ? off[t + 1] - off[t] : 0u;
    return GS3D_OK;
}

int gs3d_set_debug_mode(gs3d_ctx* c, uint32_t mode, uint32_t debug_ref) {
    if (!c) return GS3D_E_INVALID_ARG;
    if (mode > 1) return fail(c, GS3D_E_INVALID_ARG, "debug mode %u (0 = normal, 1 = tile heat map)", mode);
    c->debug_mode = mode;
    c->debug_ref = debug_ref;
    return GS3D_OK;
}

int gs3d_get_debug_mode(const gs3d_ctx* c, uint32_t* mode, uint32_t* debug_ref) {
    if (!c) return GS3D_E_INVALID_ARG;
    if (mode) *mode = c->debug_mode;
    if (debug_ref) *debug_ref = c->debug_ref;
    return GS3D_OK;
}

int gs3d_composite(gs3d_ctx* c, int format, const float* clear_rgba, uint8_t* framebuffer, size_t bytes) {
    if (!c) return GS3D_E_INVALID_ARG;
    if (format != GS3D_FORMAT_RGBA8UNORM && format != GS3D_FORMAT_BGRA8UNORM)
        return fail(c, GS3D_E_INVALID_ARG, "framebuffer format %d", format);
    if (!c->have_frame) return fail(c, GS3D_E_STATE, "no rebin yet");
    const size_t px = (size_t)c->width * c->height;
    if (framebuffer && bytes < px * 4)
        return fail(c, GS3D_E_TOO_SMALL, "framebuffer needs %zu bytes, got %zu", px * 4, bytes);
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    uint32_t* fb;
    GS3D_TRY(ensure_t(c, "framebuffer", px, &fb));
    if (!clear_rgba && framebuffer)
        HIP_TRY(c, hipMemcpyAsync(fb, framebuffer, px * 4, hipMemcpyHostToDevice, c->stream));
    CompositeParams p{};
    p.target = buf<uint16_t>(c, "splat_target");
    p.framebuffer = fb;
    p.tile_offsets = buf<uint32_t>(c, "tile_offsets");
    p.width = c->width;
    p.height = c->height;
    p.tile_x = c->tile_x;
    p.tile_y = c->tile_y;
    p.tiles_x = c->tiles_x;
    p.tiles_y = c->tiles_y;
    p.debug_ref = c->debug_ref;
    p.bgra = format == GS3D_FORMAT_BGRA8UNORM;
    p.use_clear = clear_rgba != nullptr;
    p.clear = clear_rgba ? pack_unorm8_clear(clear_rgba, p.bgra) : 0u;
    p.debug = c->debug_mode == 1;
    if (p.debug) GS3D_TRY(ensure_t(c, "debug_target", px * 4, &p.debug_target));
    // timed like a rebin stage when the frame's rebin was profiled
    const bool timed = c->last_profiled && (c->profile_mask >> GS3D_STAGE_COMPOSITE & 1u);
    if (timed) {
        GS3D_TRY(fold_composite(c));
        HIP_TRY(c, hipEventRecord(c->composite_ev[0], c->stream));
    }
    HIP_TRY(c, launch_composite(p, c->stream));
    if (timed) {
        HIP_TRY(c, hipEventRecord(c->composite_ev[1], c->stream));
        c->composite_pending = true;
    }
    c->framebuffer_px = px;
    if (p.debug) c->debug_target_px = px;
    if (framebuffer) {
        HIP_TRY(c, hipMemcpyAsync(framebuffer, fb, px * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return GS3D_OK;
}

int gs3d_stage_times_us(gs3d_ctx* c, float* out, int n) {
    if (!c || !out) return GS3D_E_INVALID_ARG;
    if (!c->have_frame || !c->last_profiled) return fail(c, GS3D_E_STATE, "no profiled rebin yet");
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    GS3D_TRY(fold_composite(c));
    for (int s = 0; s < GS3D_STAGE__COUNT && s < n; s++) {
        float ms = 0.f;
        if (c->ring_mask[c->cur_slot] >> s & 1u)
            HIP_TRY(c, hipEventElapsedTime(&ms, c->ring[c->cur_slot][s][0], c->ring[c->cur_slot][s][1]));
        out[s] = s == GS3D_STAGE_COMPOSITE ? c->composite_last_us : ms * 1000.0f;
    }
    return GS3D_OK;
}

int gs3d_profiler_reset(gs3d_ctx* c) {
    if (!c) return GS3D_E_INVALID_ARG;
    GS3D_TRY(set_device(c));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int r = 0; r < gs3d_ctx::kRing; r++) c->ring_unread[r] = false;
    c->composite_pending = false;
    c->composite_last_us = 0.f;
    for (int s = 0; s < GS3D_STAGE__COUNT; s++) {
        c->stage_sum_us[s] = 0.0;
        c->stage_count[s] = 0;
    }
    c->stage_frames = 0;
    return GS3D_OK;
}

int gs3d_set_profile_mask(gs3d_ctx* c, uint32_t stage_mask) {
    if (!c) return GS3D_E_INVALID_ARG;
    c->profile_mask = stage_mask & ((1u << GS3D_STAGE__COUNT) - 1u);
    return GS3D_OK;
}

int gs3d_stage_times_avg_us(gs3d_ctx* c, float* out, int n, uint64_t* frames) {
    if (!c || !out) return GS3D_E_INVALID_ARG;
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    for (int r = 0; r < gs3d_ctx::kRing; r++) GS3D_TRY(fold_slot(c, r));
    GS3D_TRY(fold_composite(c));
    if (frames) *frames = c->stage_frames;
    for (int s = 0; s < GS3D_STAGE__COUNT && s < n; s++)
        out[s] = c->stage_count[s] ? (float)(c->stage_sum_us[s] / (double)c->stage_count[s]) : 0.0f;
    return GS3D_OK;
}

// ---- named buffers (BufferManager, src/BufferManager.ts:32-154) -----------

int gs3d_buffer_create(gs3d_ctx* c, const char* name, size_t bytes) {
    if (!c || !name || !*name) return GS3D_E_INVALID_ARG;
    GS3D_TRY(set_device(c));
    void* p;
    GS3D_TRY(ensure(c, name, bytes < 4 ? 4 : bytes, &p));  // initBuffers: at least 4 bytes
    return GS3D_OK;
}

int gs3d_buffer_resize(gs3d_ctx* c, const char* name, size_t bytes) {
    if (!c || !name) return GS3D_E_INVALID_ARG;
    if (!c->bufs.count(name)) return fail(c, GS3D_E_UNKNOWN_NAME, "no buffer '%s'", name);
    return gs3d_buffer_create(c, name, bytes);
}

int gs3d_buffer_size(gs3d_ctx* c, const char* name, size_t* bytes) {
    if (!c || !name || !bytes) return GS3D_E_INVALID_ARG;
    auto it = c->bufs.find(name);
    if (it == c->bufs.end()) return fail(c, GS3D_E_UNKNOWN_NAME, "no buffer '%s'", name);
    *bytes = it->second.size;
    return GS3D_OK;
}

int gs3d_buffer_write(gs3d_ctx* c, const char* name, size_t offset, const void* src, size_t bytes,
                      size_t* written) {
    if (!c || !name || (bytes && !src)) return GS3D_E_INVALID_ARG;
    auto it = c->bufs.find(name);
    if (it == c->bufs.end()) return fail(c, GS3D_E_UNKNOWN_NAME, "no buffer '%s'", name);
    const size_t size = it->second.size;
    // an over-long write is clipped to the buffer; nothing left -> skipped
    const size_t n = offset < size ? std::min(bytes, size - offset) : 0;
    if (written) *written = n;
    if (n == 0) return fail(c, GS3D_E_TOO_SMALL, "write to '%s' skipped: size exceeds buffer limit", name);
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    HIP_TRY(c, hipMemcpyAsync(static_cast<char*>(it->second.ptr) + offset, src, n,
                              hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GS3D_OK;
}

int gs3d_buffer_read(gs3d_ctx* c, const char* name, size_t offset, void* dst, size_t bytes) {
    if (!c || !name || (bytes && !dst)) return GS3D_E_INVALID_ARG;
    auto it = c->bufs.find(name);
    if (it == c->bufs.end()) return fail(c, GS3D_E_UNKNOWN_NAME, "no buffer '%s'", name);
    if (offset > it->second.size || bytes > it->second.size - offset)
        return fail(c, GS3D_E_TOO_SMALL, "read of %zu bytes at %zu past '%s' (%zu bytes)", bytes, offset,
                    name, it->second.size);
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    if (bytes)
        HIP_TRY(c, hipMemcpyAsync(dst, static_cast<const char*>(it->second.ptr) + offset, bytes,
                                  hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GS3D_OK;
}

int gs3d_buffer_destroy(gs3d_ctx* c, const char* name) {
    if (!c || !name) return GS3D_E_INVALID_ARG;
    auto it = c->bufs.find(name);
    if (it == c->bufs.end()) return GS3D_OK;  // destroying an unknown name is a no-op
    GS3D_TRY(set_device(c));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (it->second.ptr) HIP_TRY(c, hipFree(it->second.ptr));
    c->bufs.erase(it);
    return GS3D_OK;
}

// Names and sizes of every buffer, largest first (Profiler.getBuffersSortedBySize,
// src/Profiler.ts:111-123). names: n_max * 64 chars; sizes: n_max entries.
int gs3d_buffer_list(gs3d_ctx* c, char* names, uint64_t* sizes, size_t n_max, size_t* count) {
    if (!c || !count) return GS3D_E_INVALID_ARG;
    std::vector<std::pair<std::string, size_t>> all;
    for (auto& kv : c->bufs) all.emplace_back(kv.first, kv.second.size);
    std::sort(all.begin(), all.end(), [](const auto& a, const auto& b) {
        return a.second != b.second ? a.second > b.second : a.first < b.first;
    });
    *count = all.size();
    for (size_t i = 0; i < all.size() && i < n_max; i++) {
        if (names) std::snprintf(names + 64 * i, 64, "%s", all[i].first.c_str());
        if (sizes) sizes[i] = all[i].second;
    }
    return GS3D_OK;
}

// ---- stage entry points --------------------------------------------------

int gs3d_stage_preprocess(gs3d_ctx* c, const gs3d_camera* cam, uint32_t tiles_x, uint32_t tiles_y,
                          float inv_tile_x, float inv_tile_y, gs3d_projected_splat* projected,
                          uint32_t* counts) {
    if (!c || !cam || (!projected && c->n) || (!counts && c->n)) return GS3D_E_INVALID_ARG;
    if (tiles_x == 0 || tiles_y == 0) return fail(c, GS3D_E_INVALID_ARG, "zero tile count");
    if (!c->bufs.count("splat_positions")) return fail(c, GS3D_E_STATE, "preprocess before set_splats");
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    const uint32_t n = c->n;
    gs3d_projected_splat* dp;
    uint32_t* dc;
    GS3D_TRY(ensure_t(c, "stage_projected", n, &dp));
    GS3D_TRY(ensure_t(c, "stage_counts", n, &dc));
    PreprocessArgs pa{};
    pa.cam = *cam;
    pa.pos = buf<float4>(c, "splat_positions");
    pa.scale = buf<float4>(c, "splat_scales");
    pa.rot = buf<float4>(c, "splat_rotations");
    pa.color = buf<float4>(c, "splat_colors");
    pa.sh = c->sh_degree > 0 ? buf<float>(c, "splat_sh") : nullptr;
    pa.out = dp;
    pa.counts = dc;
    pa.n = n;
    pa.sh_degree = c->sh_degree;
    pa.tiles_x = tiles_x;
    pa.tiles_y = tiles_y;
    pa.inv_tile_x = inv_tile_x;
    pa.inv_tile_y = inv_tile_y;
    HIP_TRY(c, launch_preprocess(pa, c->stream));
    if (n) {
        HIP_TRY(c, hipMemcpyAsync(projected, dp, (size_t)n * 64, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(counts, dc, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GS3D_OK;
}

int gs3d_stage_scan(gs3d_ctx* c, const uint32_t* counts, uint32_t n, uint32_t* offsets,
                    uint32_t* total) {
    if (!c || !offsets || (n && !counts)) return GS3D_E_INVALID_ARG;
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    uint32_t *din, *dout, *ctr;
    uint64_t* st;
    GS3D_TRY(ensure_t(c, "stage_counts", n, &din));
    GS3D_TRY(ensure_t(c, "stage_offsets", (size_t)n + 1, &dout));
    GS3D_TRY(ensure_t(c, "stage_counters", kCtrTotalWords, &ctr));
    GS3D_TRY(ensure_status(c, "stage_scan_status", (size_t)n / kScanTile + 2, &st));
    if (n) HIP_TRY(c, hipMemcpyAsync(din, counts, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, launch_fill_u32(ctr, 0u, kCtrTotalWords, c->stream));
    HIP_TRY(c, launch_scan(din, nullptr, dout, ctr + kCtrRefCounter, st, ctr + kCtrScanTicket, n,
                           next_epoch(c), c->stream));
    HIP_TRY(c, hipMemcpyAsync(offsets, dout, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost, c->stream));
    uint32_t tot = 0;
    HIP_TRY(c, hipMemcpyAsync(&tot, ctr + kCtrRefCounter, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (total) *total = tot;
    return GS3D_OK;
}

int gs3d_stage_emit(gs3d_ctx* c, const gs3d_projected_splat* projected, const uint32_t* offsets,
                    uint32_t n, uint32_t tiles_x, uint64_t* keys, uint32_t* values) {
    if (!c || !offsets || (n && !projected)) return GS3D_E_INVALID_ARG;
    if (tiles_x == 0) return fail(c, GS3D_E_INVALID_ARG, "zero tiles_x");
    const uint32_t K = offsets[n];
    if (K && (!keys || !values)) return GS3D_E_INVALID_ARG;
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    gs3d_projected_splat* dp;
    uint32_t *doff, *dv;
    uint64_t* dk;
    GS3D_TRY(ensure_t(c, "stage_projected", n, &dp));
    GS3D_TRY(ensure_t(c, "stage_offsets", (size_t)n + 1, &doff));
    GS3D_TRY(ensure_t(c, "stage_keys_a", K, &dk));
    GS3D_TRY(ensure_t(c, "stage_values_a", K, &dv));
    if (n) HIP_TRY(c, hipMemcpyAsync(dp, projected, (size_t)n * 64, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(doff, offsets, ((size_t)n + 1) * 4, hipMemcpyHostToDevice, c->stream));
    EmitArgs ea{};
    ea.projected = dp;
    ea.offsets = doff;
    ea.keys = dk;
    ea.values = dv;
    ea.hist = nullptr;
    ea.hist_partial = nullptr;
    ea.order = nullptr;
    ea.n = n;
    ea.tiles_x = tiles_x;
    ea.capacity = K;
    ea.n_passes = 0;
    ea.digit_bits = 8;
    HIP_TRY(c, launch_emit(ea, c->stream));
    if (K) {
        HIP_TRY(c, hipMemcpyAsync(keys, dk, (size_t)K * 8, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipMemcpyAsync(values, dv, (size_t)K * 4, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GS3D_OK;
}

int gs3d_stage_sort(gs3d_ctx* c, const uint64_t* keys_in, const uint32_t* values_in, uint32_t n,
                    uint32_t digit_bits, uint64_t* keys_out, uint32_t* values_out) {
    if (!c || (n && (!keys_in || !values_in || !keys_out || !values_out))) return GS3D_E_INVALID_ARG;
    if (digit_bits != 4 && digit_bits != 8)
        return fail(c, GS3D_E_INVALID_ARG, "digit width %u not 4 or 8", digit_bits);
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    if (n == 0) return GS3D_OK;
    uint64_t *ka, *kb;
    uint32_t *va, *vb, *ctr, *partial;
    GS3D_TRY(ensure_t(c, "stage_keys_a", n, &ka));
    GS3D_TRY(ensure_t(c, "stage_keys_b", n, &kb));
    GS3D_TRY(ensure_t(c, "stage_values_a", n, &va));
    GS3D_TRY(ensure_t(c, "stage_values_b", n, &vb));
    GS3D_TRY(ensure_t(c, "stage_counters", kCtrTotalWords, &ctr));
    GS3D_TRY(ensure_t(c, "stage_hist_partial", (size_t)key_hist_rows(n) * kMaxPasses * 256, &partial));
    const uint32_t passes = 64 / digit_bits;
    uint32_t shifts[kMaxPasses];
    for (uint32_t p = 0; p < passes; p++) shifts[p] = p * digit_bits;
    uint32_t host_ctr[kCtrWords] = {};
    host_ctr[kCtrRefCounter] = n;
    HIP_TRY(c, hipMemcpyAsync(ka, keys_in, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(va, values_in, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, launch_fill_u32(ctr, 0u, kCtrTotalWords, c->stream));
    HIP_TRY(c, hipMemcpyAsync(ctr, host_ctr, sizeof host_ctr, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, launch_key_histogram(ka, 8, ctr + kCtrRefCounter, n, ctr + kCtrHist, partial, digit_bits,
                                    passes, 0, c->stream));
    const uint32_t saved_bits = c->digit_bits;
    c->digit_bits = digit_bits;
    bool in_b = false;
    int rc = run_passes(c, ka, va, kb, vb, 8, false, ctr + kCtrRefCounter, n, shifts, passes,
                        ctr + kCtrHist, ctr + kCtrSortTicket, &in_b, nullptr, nullptr, 64, c->sort_key_tile);
    c->digit_bits = saved_bits;
    if (rc != GS3D_OK) return rc;
    HIP_TRY(c, hipMemcpyAsync(keys_out, in_b ? kb : ka, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(values_out, in_b ? vb : va, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GS3D_OK;
}

int gs3d_stage_identify(gs3d_ctx* c, const uint64_t* sorted_keys, uint32_t n, uint32_t tile_count,
                        uint32_t* tile_offsets) {
    if (!c || !tile_offsets || (n && !sorted_keys)) return GS3D_E_INVALID_ARG;
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    uint64_t* dk;
    uint32_t *dt, *ctr;
    GS3D_TRY(ensure_t(c, "stage_keys_a", n, &dk));
    GS3D_TRY(ensure_t(c, "stage_tile_offsets", (size_t)tile_count + 1, &dt));
    GS3D_TRY(ensure_t(c, "stage_counters", kCtrTotalWords, &ctr));
    uint32_t host_ctr[kCtrWords] = {};
    host_ctr[kCtrRefCounter] = n;
    if (n) HIP_TRY(c, hipMemcpyAsync(dk, sorted_keys, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(ctr, host_ctr, sizeof host_ctr, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, launch_identify_tiles(dk, 8, ctr + kCtrRefCounter, n, dt, tile_count, nullptr, c->stream));
    HIP_TRY(c, hipMemcpyAsync(tile_offsets, dt, ((size_t)tile_count + 1) * 4, hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GS3D_OK;
}

int gs3d_stage_rasterize(gs3d_ctx* c, const gs3d_projected_splat* projected, uint32_t n_splats,
                         const uint32_t* sorted_values, uint32_t n_refs,
                         const uint32_t* tile_offsets, uint32_t tile_x, uint32_t tile_y,
                         uint32_t width, uint32_t height, float* image) {
    if (!c || !tile_offsets || !image || (n_splats && !projected) || (n_refs && !sorted_values))
        return GS3D_E_INVALID_ARG;
    if (width == 0 || height == 0) return fail(c, GS3D_E_INVALID_ARG, "zero viewport");
    if (!((tile_x == 8 && tile_y == 8) || (tile_x == 16 && tile_y == 16) || (tile_x == 32 && tile_y == 32)))
        return fail(c, GS3D_E_UNKNOWN_NAME, "no rasterize-tiles variant for %ux%u", tile_x, tile_y);
    const uint32_t tiles_x = (width + tile_x - 1) / tile_x, tiles_y = (height + tile_y - 1) / tile_y;
    const size_t T = (size_t)tiles_x * tiles_y;
    // every range a tile will read must lie inside the caller's arrays
    for (size_t t = 0; t < T; t++) {
        const uint32_t s = tile_offsets[t], e = tile_offsets[t + 1];
        if (e > s && e > n_refs) return fail(c, GS3D_E_INVALID_ARG, "tile %zu range past n_refs", t);
    }
    for (uint32_t i = 0; i < n_refs; i++)
        if (sorted_values[i] >= n_splats) return fail(c, GS3D_E_INVALID_ARG, "splat id %u out of range", sorted_values[i]);
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    gs3d_projected_splat* dp;
    uint32_t *dv, *dt;
    uint16_t* dimg;
    float* dimg32;
    GS3D_TRY(ensure_t(c, "stage_projected", n_splats, &dp));
    GS3D_TRY(ensure_t(c, "stage_values_a", n_refs, &dv));
    GS3D_TRY(ensure_t(c, "stage_tile_offsets", T + 1, &dt));
    GS3D_TRY(ensure_t(c, "stage_image", (size_t)width * height * 4, &dimg));
    GS3D_TRY(ensure_t(c, "stage_image_f32", (size_t)width * height * 4, &dimg32));
    if (n_splats)
        HIP_TRY(c, hipMemcpyAsync(dp, projected, (size_t)n_splats * 64, hipMemcpyHostToDevice, c->stream));
    if (n_refs)
        HIP_TRY(c, hipMemcpyAsync(dv, sorted_values, (size_t)n_refs * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(dt, tile_offsets, (T + 1) * 4, hipMemcpyHostToDevice, c->stream));
    const uint32_t ppl = raster_variant_exists(tile_x, c->raster_ppl) ? c->raster_ppl : 1u;
    const bool quarters = c->raster_quarters == 2 ? quarter_lists_auto(n_refs, n_splats, tile_x, tile_y)
                                                   : c->raster_quarters != 0;
    HIP_TRY(c, launch_rasterize(dp, dv, dt, tile_x, tile_y, tiles_x, tiles_y, width, height, ppl, dimg,
                                c->stream, quarters));
    HIP_TRY(c, launch_f16_to_f32(dimg, dimg32, (size_t)width * height * 4, c->stream));
    HIP_TRY(c, hipMemcpyAsync(image, dimg32, (size_t)width * height * 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GS3D_OK;
}

int gs3d_rebin_batch(gs3d_ctx* c, const gs3d_view* views, uint32_t n_views, uint32_t flags) {
    if (!c || !views || n_views == 0) return GS3D_E_INVALID_ARG;
    if (n_views > GS3D_MAX_BATCH_VIEWS) return fail(c, GS3D_E_INVALID_ARG, "%u views (at most %u)", n_views, GS3D_MAX_BATCH_VIEWS);
    for (uint32_t i = 0; i < n_views; i++)
        if (views[i].width == 0 || views[i].height == 0 || views[i].width > 16384 || views[i].height > 16384)
            return fail(c, GS3D_E_INVALID_ARG, "view %u: viewport %ux%u", i, views[i].width, views[i].height);
    if (c->n == 0 && !c->bufs.count("splat_positions")) return fail(c, GS3D_E_STATE, "no splats uploaded");
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    return rebin_batch(c, views, n_views, flags);
}

int gs3d_batch_ref_counts(gs3d_ctx* c, uint64_t* out, uint32_t n) {
    if (!c || !out) return GS3D_E_INVALID_ARG;
    if (!c->batch || !c->batch->have) return fail(c, GS3D_E_STATE, "no batch yet");
    if (n < c->batch->K.size()) return fail(c, GS3D_E_TOO_SMALL, "need %zu counts", c->batch->K.size());
    for (size_t i = 0; i < c->batch->K.size(); i++) out[i] = c->batch->K[i];
    return GS3D_OK;
}

// A batch keeps 48-B records of its views' visible splats only: the 64-B
// gs3d_projected_splat records of one view (zeros where culled) are made
// again by the batch's projection kernel, with the batch's cameras, tiles and
// learned colours, into a readback buffer.
int batch_view_records(gs3d_ctx* c, uint32_t view, const void** out) {
    BatchState* b = c->batch;
    if (b->splat_gen != c->splat_gen)
        return fail(c, GS3D_E_STATE, "the splats changed after the batch; its records are gone");
    const BatchView& v = b->views[view];
    const uint32_t n = b->n;
    gs3d_projected_splat* rec;
    ViewParams* one;
    GS3D_TRY(ensure_t(c, "readback_projected", std::max<size_t>(n, 1), &rec));
    GS3D_TRY(ensure_t(c, "readback_view", 1, &one));
    ViewParams vp{};
    vp.cam = v.cam;
    vp.tiles_x = v.tiles_x;
    vp.tiles_y = v.tiles_y;
    vp.inv_tile_x = vp.inv_tile_y = (float)(1.0 / (double)b->tile);
    vp.out = reinterpret_cast<float4*>(rec);
    // vp lives on this stack frame: the copy completes before it returns
    HIP_TRY(c, hipMemcpyAsync(one, &vp, sizeof(vp), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    MultiPreprocessArgs ma{};
    ma.views = one;
    ma.n_views = 1;
    ma.view_colors = b->learned ? buf<uint2>(c, "batch_view_colors") + (size_t)view * n : nullptr;
    ma.pos = buf<float4>(c, "splat_positions");
    ma.scale = buf<float4>(c, "splat_scales");
    ma.rot = buf<float4>(c, "splat_rotations");
    ma.color = buf<float4>(c, "splat_colors");
    ma.sh = b->deg > 0 ? buf<float>(c, "splat_sh") : nullptr;
    ma.n = n;
    ma.sh_degree = b->deg;
    HIP_TRY(c, launch_preprocess_views_full(ma, c->stream));
    *out = rec;
    return GS3D_OK;
}

// A batch view's sorted 64-bit keys, rebuilt on the host from what the batch
// keeps: the sorted splat ids, the tile ranges (position p of tile t's range
// holds tile id t) and the view's depths (its records, projected again).
int batch_view_keys(gs3d_ctx* c, uint32_t view, uint64_t* dst, size_t* written) {
    BatchState* b = c->batch;
    const BatchView& v = b->views[view];
    const size_t K = b->K[view], T = (size_t)v.tiles_x * v.tiles_y;
    std::vector<uint32_t> vals(K), to(T + 1);
    std::vector<gs3d_projected_splat> rec(b->n);
    const void* rsrc = nullptr;
    if (b->n) GS3D_TRY(batch_view_records(c, view, &rsrc));
    if (K) HIP_TRY(c, hipMemcpyAsync(vals.data(), buf<uint32_t>(c, "batch_sorted_values") + v.val_off, K * 4,
                                     hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(to.data(), buf<uint32_t>(c, "batch_tile_offsets") + v.to_off, (T + 1) * 4,
                              hipMemcpyDeviceToHost, c->stream));
    if (b->n) HIP_TRY(c, hipMemcpyAsync(rec.data(), rsrc, (size_t)b->n * sizeof(gs3d_projected_splat),
                                        hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    size_t filled = 0;
    for (size_t t = 0; t < T; t++) {
        const size_t lo = to[t], hi = to[t + 1];
        for (size_t p = lo; p < hi && p < K; p++) {
            float depth = rec[vals[p]].depth;
            depth = -depth;
            uint32_t bits;
            std::memcpy(&bits, &depth, 4);
            dst[p] = ((uint64_t)t << 32) | bits;
            filled++;
        }
    }
    if (filled != K) return fail(c, GS3D_E_STATE, "tile ranges of view %u cover %zu of %zu keys", view, filled, K);
    if (written) *written = K * 8;
    return GS3D_OK;
}

int gs3d_batch_read(gs3d_ctx* c, uint32_t view, int id, void* dst, size_t bytes, size_t* written) {
    if (!c) return GS3D_E_INVALID_ARG;
    BatchState* b = c->batch;
    if (!b || !b->have) return fail(c, GS3D_E_STATE, "no batch yet");
    if (view >= b->views.size()) return fail(c, GS3D_E_INVALID_ARG, "view %u of %zu", view, b->views.size());
    GS3D_TRY(set_device(c));
    const BatchView& v = b->views[view];
    const size_t px = (size_t)v.W * v.H;
    const void* src = nullptr;
    size_t size = 0;
    bool widen = false;
    switch (id) {
        case GS3D_BUF_IMAGE_RGBA16F: src = buf<uint16_t>(c, "batch_images") + v.img_off; size = px * 8; break;
        case GS3D_BUF_IMAGE_RGBA32F:
            src = buf<uint16_t>(c, "batch_images") + v.img_off; size = px * 16; widen = true; break;
        case GS3D_BUF_TILE_OFFSETS:
            src = buf<uint32_t>(c, "batch_tile_offsets") + v.to_off; size = ((size_t)v.tiles_x * v.tiles_y + 1) * 4; break;
        case GS3D_BUF_PROJECTED_SPLATS:
            size = (size_t)b->n * sizeof(gs3d_projected_splat);
            if (dst && bytes >= size) GS3D_TRY(batch_view_records(c, view, &src));
            break;
        case GS3D_BUF_REF_COUNTER:
            src = buf<uint32_t>(c, "batch_counters") + (size_t)view * kCtrTotalWords + kCtrRefCounter; size = 4; break;
        case GS3D_BUF_SORT_VALUES:
            src = buf<uint32_t>(c, "batch_sorted_values") + v.val_off; size = (size_t)b->K[view] * 4; break;
        case GS3D_BUF_SORT_KEYS:
            size = (size_t)b->K[view] * 8;
            if (dst && bytes >= size) return batch_view_keys(c, view, static_cast<uint64_t*>(dst), written);
            break;
        default: return fail(c, GS3D_E_UNKNOWN_NAME, "buffer %d is not kept per batch view", id);
    }
    if (written) *written = size;
    if (!dst) return GS3D_OK;
    if (bytes < size) return fail(c, GS3D_E_TOO_SMALL, "buffer %d needs %zu bytes, got %zu", id, size, bytes);
    if (size == 0) return GS3D_OK;
    if (widen) {
        float* tmp;
        GS3D_TRY(ensure_t(c, "readback_f32", px * 4, &tmp));
        HIP_TRY(c, launch_f16_to_f32(static_cast<const uint16_t*>(src), tmp, px * 4, c->stream));
        src = tmp;
    }
    HIP_TRY(c, hipMemcpyAsync(dst, src, size, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GS3D_OK;
}

int gs3d_learned_dirs(float* dirs) {
    if (!dirs) return GS3D_E_INVALID_ARG;
    learned_directions(dirs);
    return GS3D_OK;
}

int gs3d_learned_init(gs3d_ctx* c, uint64_t seed) {
    if (!c) return GS3D_E_INVALID_ARG;
    if (!c->bufs.count("splat_positions")) return fail(c, GS3D_E_STATE, "no splats uploaded");
    if (c->sh_degree == 0) return fail(c, GS3D_E_STATE, "learned colour needs SH coefficients (degree >= 1)");
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    c->learned_n = c->n;
    GS3D_TRY(learned_buffers(c));
    uint64_t st = seed * 0x2545F4914F6CDD1Dull + 1;
    std::vector<float> p(kLearnedParams, 0.0f);
    const double s1 = std::sqrt(2.0 / 11.0), s2 = std::sqrt(2.0 / 64.0), s3 = std::sqrt(1.0 / 64.0);
    for (int i = 0; i < 64 * 11; i++) p[i] = (float)(s1 * normal01(st));
    for (int i = 0; i < 64 * 64; i++) p[64 * 11 + 64 + i] = (float)(s2 * normal01(st));
    for (int i = 0; i < 3 * 64; i++) p[64 * 11 + 64 + 64 * 64 + 64 + i] = (float)(s3 * normal01(st));
    std::vector<float> z((size_t)c->n * 8);
    for (auto& v : z) v = (float)(0.1 * normal01(st));
    std::vector<float> dirs(GS3D_LEARNED_DIRS * 3), basis(GS3D_LEARNED_DIRS * 16);
    learned_directions(dirs.data());
    for (int k = 0; k < GS3D_LEARNED_DIRS; k++) {
        double b[16];
        sh_basis_f64(dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2], b);
        for (int i = 0; i < 16; i++) basis[16 * k + i] = (float)b[i];
    }
    HIP_TRY(c, hipMemcpyAsync(buf<void>(c, "learned_params"), p.data(), p.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(buf<void>(c, "learned_latents"), z.data(), z.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(buf<void>(c, "learned_dirs"), dirs.data(), dirs.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(buf<void>(c, "learned_basis"), basis.data(), basis.size() * 4, hipMemcpyHostToDevice, c->stream));
    for (const char* name : {"learned_pm", "learned_pv"})
        HIP_TRY(c, launch_fill_u32(buf<uint32_t>(c, name), 0u, kLearnedParams, c->stream));
    for (const char* name : {"learned_lm", "learned_lv"})
        HIP_TRY(c, launch_fill_u32(buf<uint32_t>(c, name), 0u, (size_t)c->n * 8, c->stream));
    HIP_TRY(c, launch_f16_copy(buf<float>(c, "learned_latents"), (size_t)c->n * 8,
                               buf<uint16_t>(c, "learned_latents_f16"), c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->learned_init = true;
    c->learned_enabled = false;
    c->learned_step = 0;
    return GS3D_OK;
}

int gs3d_learned_train(gs3d_ctx* c, uint32_t epochs, float lr, float lr_latent, float* loss, uint32_t n_loss) {
    if (!c) return GS3D_E_INVALID_ARG;
    GS3D_TRY(learned_require(c));
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    const size_t n8 = (size_t)c->n * 8;
    for (uint32_t e = 0; e < epochs; e++) {
        for (uint32_t s = 0; s < GS3D_LEARNED_DIRS / 16; s++) {
            GS3D_TRY(learned_step_grads(c, 16 * s, s));
            c->learned_step++;
            HIP_TRY(c, launch_adam(buf<float>(c, "learned_params"), buf<float>(c, "learned_pm"),
                                   buf<float>(c, "learned_pv"), buf<float>(c, "learned_grad"), kLearnedParams, lr,
                                   0.9f, 0.999f, 1e-8f, c->learned_step, nullptr, c->stream));
            HIP_TRY(c, launch_adam(buf<float>(c, "learned_latents"), buf<float>(c, "learned_lm"),
                                   buf<float>(c, "learned_lv"), buf<float>(c, "learned_lgrad"), n8, lr_latent, 0.9f,
                                   0.999f, 1e-8f, c->learned_step, buf<uint16_t>(c, "learned_latents_f16"), c->stream));
        }
        if (loss && e < n_loss) {
            float l[16];
            HIP_TRY(c, hipMemcpyAsync(l, buf<float>(c, "learned_loss"), sizeof l, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            double t = 0;
            for (float v : l) t += v;
            loss[e] = (float)(t / (3.0 * GS3D_LEARNED_DIRS * std::max(1u, c->n)));
        }
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->learned_enabled = false;  // parameters changed: validate again
    return GS3D_OK;
}

int gs3d_learned_grad(gs3d_ctx* c, uint32_t first_dir, float* grad_params, float* grad_latents, float* loss) {
    if (!c) return GS3D_E_INVALID_ARG;
    if (first_dir % 16 || first_dir + 16 > GS3D_LEARNED_DIRS) return fail(c, GS3D_E_INVALID_ARG, "first_dir %u", first_dir);
    GS3D_TRY(learned_require(c));
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    GS3D_TRY(learned_step_grads(c, first_dir, 0));
    float l = 0.f;
    if (grad_params)
        HIP_TRY(c, hipMemcpyAsync(grad_params, buf<float>(c, "learned_grad"), kLearnedParams * 4, hipMemcpyDeviceToHost, c->stream));
    if (grad_latents)
        HIP_TRY(c, hipMemcpyAsync(grad_latents, buf<float>(c, "learned_lgrad"), (size_t)c->n * 32, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipMemcpyAsync(&l, buf<float>(c, "learned_loss"), 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = l / (3.0f * 16.0f * (float)std::max(1u, c->n));
    return GS3D_OK;
}

int gs3d_learned_get(gs3d_ctx* c, float* params, float* latents, uint16_t* latents_f16) {
    if (!c) return GS3D_E_INVALID_ARG;
    GS3D_TRY(learned_require(c));
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    if (params) HIP_TRY(c, hipMemcpyAsync(params, buf<float>(c, "learned_params"), kLearnedParams * 4, hipMemcpyDeviceToHost, c->stream));
    if (latents) HIP_TRY(c, hipMemcpyAsync(latents, buf<float>(c, "learned_latents"), (size_t)c->n * 32, hipMemcpyDeviceToHost, c->stream));
    if (latents_f16) HIP_TRY(c, hipMemcpyAsync(latents_f16, buf<uint16_t>(c, "learned_latents_f16"), (size_t)c->n * 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GS3D_OK;
}

int gs3d_learned_set(gs3d_ctx* c, const float* params, const float* latents) {
    if (!c) return GS3D_E_INVALID_ARG;
    GS3D_TRY(learned_require(c));
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    if (params) HIP_TRY(c, hipMemcpyAsync(buf<float>(c, "learned_params"), params, kLearnedParams * 4, hipMemcpyHostToDevice, c->stream));
    if (latents) {
        HIP_TRY(c, hipMemcpyAsync(buf<float>(c, "learned_latents"), latents, (size_t)c->n * 32, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, launch_f16_copy(buf<float>(c, "learned_latents"), (size_t)c->n * 8,
                                   buf<uint16_t>(c, "learned_latents_f16"), c->stream));
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->learned_enabled = false;
    return GS3D_OK;
}

int gs3d_learned_get_state(gs3d_ctx* c, gs3d_learned_state* st) {
    if (!c || !st) return GS3D_E_INVALID_ARG;
    GS3D_TRY(learned_require(c));
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    const size_t np = kLearnedParams * 4, nz = (size_t)c->n * 32;
    const std::pair<float*, const char*> params[] = {{st->params, "learned_params"}, {st->params_m, "learned_pm"},
                                                     {st->params_v, "learned_pv"}};
    const std::pair<float*, const char*> lat[] = {{st->latents, "learned_latents"}, {st->latents_m, "learned_lm"},
                                                  {st->latents_v, "learned_lv"}};
    for (const auto& p : params)
        if (p.first) HIP_TRY(c, hipMemcpyAsync(p.first, buf<float>(c, p.second), np, hipMemcpyDeviceToHost, c->stream));
    for (const auto& p : lat)
        if (p.first && nz) HIP_TRY(c, hipMemcpyAsync(p.first, buf<float>(c, p.second), nz, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    st->step = c->learned_step;
    return GS3D_OK;
}

int gs3d_learned_set_state(gs3d_ctx* c, const gs3d_learned_state* st) {
    if (!c || !st) return GS3D_E_INVALID_ARG;
    GS3D_TRY(learned_require(c));
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    const size_t np = kLearnedParams * 4, nz = (size_t)c->n * 32;
    const std::pair<const float*, const char*> params[] = {{st->params, "learned_params"}, {st->params_m, "learned_pm"},
                                                           {st->params_v, "learned_pv"}};
    const std::pair<const float*, const char*> lat[] = {{st->latents, "learned_latents"}, {st->latents_m, "learned_lm"},
                                                        {st->latents_v, "learned_lv"}};
    for (const auto& p : params)
        if (p.first) HIP_TRY(c, hipMemcpyAsync(buf<float>(c, p.second), p.first, np, hipMemcpyHostToDevice, c->stream));
    for (const auto& p : lat)
        if (p.first && nz) HIP_TRY(c, hipMemcpyAsync(buf<float>(c, p.second), p.first, nz, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, launch_f16_copy(buf<float>(c, "learned_latents"), (size_t)c->n * 8,
                               buf<uint16_t>(c, "learned_latents_f16"), c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->learned_step = st->step;
    c->learned_enabled = false;
    return GS3D_OK;
}

int gs3d_learned_colors(gs3d_ctx* c, const float eye[3], float* rgb) {
    if (!c || !eye || !rgb) return GS3D_E_INVALID_ARG;
    GS3D_TRY(learned_require(c));
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    float* eyes;
    uint2* colors;
    GS3D_TRY(ensure_t(c, "learned_probe_eye", 4, &eyes));
    GS3D_TRY(ensure_t(c, "learned_probe_colors", c->n, &colors));
    const float e4[4] = {eye[0], eye[1], eye[2], 1.0f};
    HIP_TRY(c, hipMemcpyAsync(eyes, e4, 16, hipMemcpyHostToDevice, c->stream));
    LearnedColorArgs la{};
    la.params = buf<float>(c, "learned_params");
    la.latents = buf<uint16_t>(c, "learned_latents_f16");
    la.pos = buf<float4>(c, "splat_positions");
    la.eyes = eyes;
    la.colors = colors;
    la.n = c->n;
    la.n_views = 1;
    HIP_TRY(c, launch_learned_colors(la, c->stream));
    std::vector<uint16_t> h((size_t)c->n * 4);
    HIP_TRY(c, hipMemcpyAsync(h.data(), colors, h.size() * 2, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < c->n; i++)
        for (int k = 0; k < 3; k++) rgb[i * 3 + k] = half_to_float_host(h[i * 4 + k]);
    return GS3D_OK;
}

int gs3d_learned_enable(gs3d_ctx* c, int on) {
    if (!c) return GS3D_E_INVALID_ARG;
    if (on) GS3D_TRY(learned_require(c));
    c->learned_enabled = on != 0;
    return GS3D_OK;
}

int gs3d_learned_validate(gs3d_ctx* c, const gs3d_view* views, uint32_t n_views, float min_psnr_db, float* psnr,
                          int* enabled) {
    if (!c || !views || n_views == 0) return GS3D_E_INVALID_ARG;
    GS3D_TRY(learned_require(c));
    GS3D_TRY(gs3d_rebin_batch(c, views, n_views, 0));
    BatchState* b = c->batch;
    size_t img_elems = 0;
    for (const auto& v : b->views) img_elems = std::max(img_elems, v.img_off + (size_t)v.W * v.H * 4);
    uint16_t* ref;
    GS3D_TRY(ensure_t(c, "learned_ref_images", img_elems, &ref));
    HIP_TRY(c, hipMemcpyAsync(ref, buf<uint16_t>(c, "batch_images"), img_elems * 2, hipMemcpyDeviceToDevice, c->stream));
    c->learned_force = true;
    const int rc = gs3d_rebin_batch(c, views, n_views, GS3D_BATCH_LEARNED);
    c->learned_force = false;
    GS3D_TRY(rc);
    double* acc;
    GS3D_TRY(ensure_t(c, "learned_sq_sum", 1, &acc));
    HIP_TRY(c, launch_fill_u32(reinterpret_cast<uint32_t*>(acc), 0u, 2, c->stream));
    size_t px = 0;
    for (const auto& v : b->views) {
        HIP_TRY(c, launch_sq_diff_sum(ref + v.img_off, buf<uint16_t>(c, "batch_images") + v.img_off,
                                      (size_t)v.W * v.H, acc, c->stream));
        px += (size_t)v.W * v.H;
    }
    double sq = 0;
    HIP_TRY(c, hipMemcpyAsync(&sq, acc, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const double mse = sq / (3.0 * (double)std::max<size_t>(1, px));
    const double p = mse > 0 ? 10.0 * std::log10(1.0 / mse) : 200.0;
    if (psnr) *psnr = (float)p;
    c->learned_enabled = p >= (double)min_psnr_db;
    if (enabled) *enabled = c->learned_enabled ? 1 : 0;
    return GS3D_OK;
}

int gs3d_stage_composite(gs3d_ctx* c, const uint16_t* target_f16, const uint32_t* tile_offsets,
                         uint32_t tile_x, uint32_t tile_y, uint32_t width, uint32_t height,
                         uint32_t debug, uint32_t debug_ref, int format, const float* clear_rgba,
                         uint8_t* framebuffer, uint16_t* debug_out) {
    if (!c || !target_f16 || !framebuffer || (debug && !tile_offsets)) return GS3D_E_INVALID_ARG;
    if (width == 0 || height == 0 || tile_x == 0 || tile_y == 0)
        return fail(c, GS3D_E_INVALID_ARG, "zero viewport or tile size");
    if ((uint64_t)width * height > (1ull << 31)) return fail(c, GS3D_E_INVALID_ARG, "viewport too large");
    if (debug > 1) return fail(c, GS3D_E_INVALID_ARG, "debug mode %u", debug);
    if (format != GS3D_FORMAT_RGBA8UNORM && format != GS3D_FORMAT_BGRA8UNORM)
        return fail(c, GS3D_E_INVALID_ARG, "framebuffer format %d", format);
    GS3D_TRY(set_device(c));
    GS3D_TRY(settle(c));
    const uint32_t tiles_x = (width + tile_x - 1) / tile_x, tiles_y = (height + tile_y - 1) / tile_y;
    const size_t px = (size_t)width * height, T = (size_t)tiles_x * tiles_y;
    uint16_t *dt, *dd = nullptr;
    uint32_t *dfb, *dto = nullptr;
    GS3D_TRY(ensure_t(c, "stage_image", px * 4, &dt));
    GS3D_TRY(ensure_t(c, "stage_framebuffer", px, &dfb));
    HIP_TRY(c, hipMemcpyAsync(dt, target_f16, px * 8, hipMemcpyHostToDevice, c->stream));
    if (!clear_rgba) HIP_TRY(c, hipMemcpyAsync(dfb, framebuffer, px * 4, hipMemcpyHostToDevice, c->stream));
    if (debug) {
        GS3D_TRY(ensure_t(c, "stage_tile_offsets", T + 1, &dto));
        GS3D_TRY(ensure_t(c, "stage_debug_target", px * 4, &dd));
        HIP_TRY(c, hipMemcpyAsync(dto, tile_offsets, (T + 1) * 4, hipMemcpyHostToDevice, c->stream));
    }
    CompositeParams p{};
    p.target = dt;
    p.debug_target = dd;
    p.framebuffer = dfb;
    p.tile_offsets = dto;
    p.width = width;
    p.height = height;
    p.tile_x = tile_x;
    p.tile_y = tile_y;
    p.tiles_x = tiles_x;
    p.tiles_y = tiles_y;
    p.debug_ref = debug_ref;
    p.bgra = format == GS3D_FORMAT_BGRA8UNORM;
    p.use_clear = clear_rgba != nullptr;
    p.clear = clear_rgba ? pack_unorm8_clear(clear_rgba, p.bgra) : 0u;
    p.debug = (int)debug;
    HIP_TRY(c, launch_composite(p, c->stream));
    HIP_TRY(c, hipMemcpyAsync(framebuffer, dfb, px * 4, hipMemcpyDeviceToHost, c->stream));
    if (debug && debug_out)
        HIP_TRY(c, hipMemcpyAsync(debug_out, dd, px * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return GS3D_OK;
}

int gs3d_sort_pairs_noalloc(gs3d_ctx* c, const uint64_t* keys_in, const uint32_t* values_in, uint32_t n,
                            uint32_t digit_bits, uint64_t* keys_out, uint32_t* values_out) {
    if (!c) return GS3D_E_INVALID_ARG;
    if (n && (!keys_in || !values_in || !keys_out || !values_out))
        return fail(c, GS3D_E_INVALID_ARG, "null device pointer");
    if (digit_bits != 4 && digit_bits != 8)
        return fail(c, GS3D_E_INVALID_ARG, "digit width %u not 4 or 8", digit_bits);
    if (n > GS3D_SORT_NOALLOC_MAX_PAIRS)
        return fail(c, GS3D_E_INVALID_ARG,
                    "%u pairs exceed the %u the allocation-free sort's static scratch holds "
                    "(gs3d_stage_sort has no such limit)", n, GS3D_SORT_NOALLOC_MAX_PAIRS);
    GS3D_TRY(set_device(c));
    uint8_t* base = noalloc_scratch(c->device);
    if (!base) return fail(c, GS3D_E_STATE, "no scratch address for device %d", c->device);
    if (n == 0) return GS3D_OK;
    uint32_t* ctl = reinterpret_cast<uint32_t*>(base);
    uint32_t* hist = reinterpret_cast<uint32_t*>(base + kNaHistOff);
    uint32_t* partial = reinterpret_cast<uint32_t*>(base + kNaPartialOff);
    uint64_t* status = reinterpret_cast<uint64_t*>(base + kNaStatusOff);
    uint64_t* tmp_keys = reinterpret_cast<uint64_t*>(base + kNaKeysOff);
    uint32_t* tmp_values = reinterpret_cast<uint32_t*>(base + kNaValuesOff);
    const uint32_t passes = 64u / digit_bits, radix = 1u << digit_bits;
    // the scratch keeps the last call's contents: control words and histograms
    // start at zero (fill kernels, no memset nodes), then the key count
    HIP_TRY(c, launch_fill_u32(ctl, 0u, kNaCtlWords, c->stream));
    HIP_TRY(c, launch_fill_u32(ctl, n, 1, c->stream));
    HIP_TRY(c, launch_fill_u32(hist, 0u, (size_t)passes * radix, c->stream));
    HIP_TRY(c, launch_key_histogram(keys_in, 8, ctl, n, hist, partial, digit_bits, passes, 0, c->stream));
    // Ping-pong between the caller's output and the scratch side so that the
    // last pass lands in the output: pass p writes the output when
    // passes - 1 - p is even. The passes are even in number, so pass 0 writes
    // the scratch side and the input may be the output buffer itself.
    const void* src_k = keys_in;
    const uint32_t* src_v = values_in;
    for (uint32_t p = 0; p < passes; p++) {
        const bool to_out = ((passes - 1u - p) % 2u) == 0u;
        void* dst_k = to_out ? static_cast<void*>(keys_out) : static_cast<void*>(tmp_keys);
        uint32_t* dst_v = to_out ? values_out : tmp_values;
        HIP_TRY(c, launch_radix_pass(src_k, src_v, dst_k, dst_v, 8, false, ctl, n, hist + p * radix, status,
                                     ctl + 1 + p, p * digit_bits, digit_bits, next_epoch(c), nullptr, nullptr,
                                     c->stream, nullptr, c->sort_key_tile, digit_bits, false));
        src_k = dst_k;
        src_v = dst_v;
    }
    return GS3D_OK;
}

}  // extern "C"
