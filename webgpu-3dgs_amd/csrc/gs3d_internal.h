// gs3d_internal.h — shared definitions of the HIP kernels and the runtime.
//
// Device memory layout (all buffers owned by a gs3d_ctx, see runtime.cpp):
//   splat_positions/scales/rotations/colors  N x float4 (the loader's layouts)
//   splat_sh          SoA f32, coefficient-major: sh[(k*3 + ch)*N + i],
//                     k = 0 (f_dc) .. (deg+1)^2-1; one coalesced load per
//                     (thread, coefficient)
//   projected_splats  N x 64 B gs3d_projected_splat
//   splat_ref_counts  N x u32 (id order)
//   splat_tile_rects  N x uint2 tile rectangle (x0 | y0 << 16, x1 | y1 << 16)
//   depth keys/order  N x u32 each, a and b (the depth presort)
//   splat_emit_records (N+1) x uint4 (first key, first tile, width, depth key)
//   sort_keys_a/b     cap x u64 (tile << 32 | bits(-depth)); values cap x u32
//   tile_offsets      (T+1) x u32
//   splat_target      W*H x 4 x f16 (premultiplied colour, 1 - T)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gs3d.h"

namespace gs3d {

constexpr int kWave = 64;

// Lookback status words (scan and radix passes): [epoch:30][flag:2][value:32].
// A word whose epoch differs from the launch's epoch reads as "not ready", so
// the status arrays never need clearing between launches.
constexpr uint64_t kFlagAggregate = 1ull;
constexpr uint64_t kFlagPrefix = 2ull;
__host__ __device__ inline uint64_t status_word(uint32_t epoch, uint64_t flag, uint32_t value) {
    return ((uint64_t)epoch << 34) | (flag << 32) | (uint64_t)value;
}
// A kernel replayed from a captured graph cannot take a fresh epoch as an
// argument. It adds a device-resident base instead, which the graph's first
// node raises by kEpochStride on every replay; the argument is then a distinct
// offset in [1, kEpochStride) per use of a status array within one replay, so
// no status word left by the previous replay can match.
// Every batch within GS3D_MAX_BATCH_VIEWS fits one stride: a view takes at
// most 32/4 depth passes + 1 scan + 8 tile-id passes = 17 epochs, and even all
// 1,024 views on one stream would take 17,408 < kEpochStride. The base wraps
// after 2^30 / kEpochStride bumps; the host zeroes the status arrays and the
// base before that, so a word of an earlier round of bases never matches.
constexpr uint32_t kEpochMask = (1u << 30) - 1u;
constexpr uint32_t kEpochStride = 32768;
constexpr uint32_t kMaxEpochsPerView = 32 / 4 + 1 + 32 / 4;
static_assert(GS3D_MAX_BATCH_VIEWS * kMaxEpochsPerView < kEpochStride, "replay epochs of a batch");
constexpr uint32_t kEpochBumpsPerRound = (kEpochMask + 1u) / kEpochStride;
__device__ __forceinline__ uint32_t resolve_epoch(uint32_t epoch, const uint32_t* epoch_base) {
    return epoch_base ? ((epoch + *epoch_base) & kEpochMask) : epoch;
}

// Scan: 1024 threads x 8 counts per thread.
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// Radix: the smallest tile (keys per workgroup) any pass shape uses; it sizes
// the lookback status words and the key-capacity slack.
constexpr int kSortTile = 4096;
constexpr int kMaxPasses = 16;

// Radix scatter destination. The keys of digit d in a tile are at tile
// positions [local_start, local_start + c); they go to global_start + excl +
// (j - local_start) for position j, where global_start is the digit's start
// over all keys and excl the count in the preceding tiles. The per-digit delta
// global_start + excl - local_start is kept in u32 and may wrap below zero;
// delta + j modulo 2^32 is then the exact index, which is below the key count
// (at most 2^32 - 1), so no index is negative or truncated for any K.
__host__ __device__ inline uint32_t scatter_delta(uint32_t global_start, uint32_t excl, uint32_t local_start) {
    return global_start + excl - local_start;
}
__host__ __device__ inline size_t scatter_index(uint32_t delta, uint32_t j) {
    return (size_t)(uint32_t)(delta + j);
}

// Counters buffer slots (u32), cleared at the start of every rebin.
constexpr int kCtrScanTicket = 0;
constexpr int kCtrSortTicket = 1;   // + pass index (up to kMaxPasses)
constexpr int kCtrWords = 64;

struct PreprocessArgs {
    gs3d_camera cam;
    const float4* pos;
    const float4* scale;
    const float4* rot;
    const float4* color;
    const float* sh;  // SoA, see above; null when sh_degree == 0
    gs3d_projected_splat* out;
    uint32_t* counts;
    uint32_t* depth_keys;  // bits(-depth) of visible splats, 0xFFFFFFFF for culled; or null
    uint2* rects;          // tile rectangle (x0 | y0 << 16, x1 | y1 << 16), zero for culled; or null
    uint32_t* rects8;      // the same rectangle as x0 | y0 << 8 | x1 << 16 | y1 << 24 when both tile
                           // counts are <= 256 (zero for culled); or null
    uint32_t* counters;    // if set: counters[0..counter_words) = 0 except [count_slot] = n
    uint32_t counter_words;
    uint32_t count_slot;
    uint32_t n;
    int sh_degree;
    uint32_t tiles_x, tiles_y;
    float inv_tile_x, inv_tile_y;
};

// One view of a batch, as the multi-view preprocess sees it (device memory).
struct ViewParams {
    gs3d_camera cam;
    uint32_t tiles_x, tiles_y;
    float inv_tile_x, inv_tile_y;
    // 48-B records (the first three float4s of a gs3d_projected_splat) of the
    // visible splats; culled splats' records are not written. The full-record
    // kernel writes gs3d_projected_splat records of every splat here instead.
    float4* out;
    uint32_t* depth_keys;
    uint2* rects;        // 8-B tile rectangles, or null
    uint32_t* rects8;    // or the packed rectangles (both tile counts <= 256), or null
    uint32_t* counters;  // counters[0..counter_words) = 0 except [count_slot] = n
};

struct MultiPreprocessArgs {
    const ViewParams* views;  // device array of n_views
    uint32_t n_views;
    const uint2* view_colors;  // null, or learned colours [n_views][n] (f16 r, g, b, 1) used instead of SH
    const float4* pos;
    const float4* scale;
    const float4* rot;
    const float4* color;
    const float* sh;
    uint32_t n;
    int sh_degree;
    uint32_t counter_words;
    uint32_t count_slot;
};
hipError_t launch_preprocess_views(const MultiPreprocessArgs& a, hipStream_t s);
// The same arithmetic writing, for every splat, the 64-B gs3d_projected_splat
// a single-view frame writes (zeros when culled): how a batch view's records
// are read back.
hipError_t launch_preprocess_views_full(const MultiPreprocessArgs& a, hipStream_t s);

struct EmitArgs {
    const gs3d_projected_splat* projected;
    const uint32_t* order;    // position -> splat id (null: identity)
    const uint32_t* offsets;  // n + 1 entries, per position (when recs is null)
    const uint4* recs;        // n + 1 records of k_scan_rects, or null
    uint64_t* keys;           // (tile << 32 | depth key) per key, or null when tiles is set
    uint32_t* tiles;          // or the tile id alone (the depth order is already in the stream)
    uint16_t* tiles16;        // or the tile id as 16 bits (frames of at most 65,536 tiles)
    uint32_t* values;
    uint32_t* hist;           // [passes][radix] digit histograms (accumulated), or null
    uint32_t* hist_partial;   // [emit_grid(n)][passes][radix] scratch rows, or null
    uint32_t n;
    uint32_t tiles_x;
    uint64_t capacity;        // keys/values entries allocated
    uint32_t n_passes;        // histogram of pass p: digit at bit hist_shift0 + p * digit_bits
    uint32_t digit_bits;
    uint32_t hist_shift0;
};

// Emit runs at most this many workgroups (grid-stride over 256-splat chunks),
// which bounds the per-workgroup histogram rows to fold.
constexpr uint32_t kEmitMaxBlocks = 1024;
uint32_t emit_grid(uint32_t n);

// Exclusive scan of the tile counts of the splats in depth order, counted from
// their tile rectangles: recs[p] = (first key, first tile id, width in tiles,
// depth key) for depth position p, recs[n].x = K = ref_counter[0].
struct ScanRectsArgs {
    const uint32_t* order;       // depth position -> splat id
    const uint32_t* depth_keys;  // depth position -> depth key, 0xFFFFFFFF for a culled splat
    const uint2* rects;          // per splat id: tile rectangle (x0 | y0 << 16, x1 | y1 << 16)
    const uint32_t* rects8;      // or, when set: per depth position, the rectangle packed 8 bits per
                                 // coordinate (carried into depth order by the last depth pass)
    uint4* recs;                 // n + 1 records
    uint32_t* ref_counter;
    uint64_t* status;            // lookback words, n / kScanTile + 2
    uint32_t* ticket;
    uint32_t n;
    uint32_t tiles_x;
    uint32_t epoch;
    const uint32_t* epoch_base;  // null, or the replay base added to epoch
};
hipError_t launch_scan_rects(const ScanRectsArgs& a, hipStream_t s);

// Lanes of the wavefront whose BITS-bit digit equals this lane's, among lanes
// with `valid` set (wave-level match by one ballot per digit bit). Per bit:
// the bit as an all-ones/zero mask m, its ballot, and lo &= ~(ballot ^ m) on
// each 32-bit half as one v_bitop3 (truth table 0x90 for src0 & ~(src1 ^
// src2), with src0, src1, src2 standing for 0xF0, 0xCC, 0xAA). Bits at or
// above match_bits (uniform) are zero in every digit and are skipped.
// Bit b of d as an all-ones/zero mask. The empty asm hides the mask's origin
// from the optimizer, so its ballot is one compare of the mask with zero
// instead of a shift of d and a compare beside the bit extract.
__device__ __forceinline__ uint32_t digit_bit_mask(uint32_t d, int b) {
    uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int32_t)d, b, 1);
    asm("" : "+v"(m));
    return m;
}

template <int BITS>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid, uint32_t match_bits = BITS) {
    const uint64_t vm = __ballot(valid);
    const uint32_t mb = __builtin_amdgcn_readfirstlane(match_bits);
    uint32_t lo = (uint32_t)vm, hi = (uint32_t)(vm >> 32);
#pragma unroll
    for (int b = 0; b < BITS; b++) {
        if ((uint32_t)b < mb) {
            const uint32_t m = digit_bit_mask(d, b);  // all ones if bit b set
            const uint64_t bal = __ballot(m != 0u);
            lo = __builtin_amdgcn_bitop3_b32(lo, (uint32_t)bal, m, 0x90);
            hi = __builtin_amdgcn_bitop3_b32(hi, (uint32_t)(bal >> 32), m, 0x90);
        }
    }
    return ((uint64_t)hi << 32) | lo;
}

// match_digit for two items at once: their ballot steps interleave, so each
// step's ballot result is read two instructions after it is written (no
// wait states) and the uniform test of a bit against match_bits is taken once
// per pair. FULL: every digit bit can be set (no test at all).
template <int BITS, bool FULL>
__device__ __forceinline__ void match_digit2(uint32_t d0, bool valid0, uint32_t d1, bool valid1, uint32_t match_bits,
                                             uint64_t& peers0, uint64_t& peers1) {
    const uint64_t vm0 = __ballot(valid0), vm1 = __ballot(valid1);
    uint32_t lo0 = (uint32_t)vm0, hi0 = (uint32_t)(vm0 >> 32);
    uint32_t lo1 = (uint32_t)vm1, hi1 = (uint32_t)(vm1 >> 32);
#pragma unroll
    for (int b = 0; b < BITS; b++) {
        if (FULL || (uint32_t)b < match_bits) {
            const uint32_t m0 = digit_bit_mask(d0, b);  // all ones if bit b set
            const uint32_t m1 = digit_bit_mask(d1, b);
            const uint64_t bal0 = __ballot(m0 != 0u);
            const uint64_t bal1 = __ballot(m1 != 0u);
            lo0 = __builtin_amdgcn_bitop3_b32(lo0, (uint32_t)bal0, m0, 0x90);
            hi0 = __builtin_amdgcn_bitop3_b32(hi0, (uint32_t)(bal0 >> 32), m0, 0x90);
            lo1 = __builtin_amdgcn_bitop3_b32(lo1, (uint32_t)bal1, m1, 0x90);
            hi1 = __builtin_amdgcn_bitop3_b32(hi1, (uint32_t)(bal1 >> 32), m1, 0x90);
        }
    }
    peers0 = ((uint64_t)hi0 << 32) | lo0;
    peers1 = ((uint64_t)hi1 << 32) | lo1;
}

__device__ __forceinline__ uint64_t match_digit_rt(uint32_t d, bool valid, uint32_t bits) {
    return bits == 8 ? match_digit<8>(d, valid) : match_digit<4>(d, valid);
}

// ---- launchers (each enqueues on `stream`, returns the HIP status) --------
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s);
hipError_t launch_sh_relayout(const float* f_dc, const float* rest, float* sh, uint32_t n, int deg,
                              hipStream_t s);
hipError_t launch_preprocess(const PreprocessArgs& a, hipStream_t s);
// Exclusive scan of counts[order[i]] (order null: counts[i]).
hipError_t launch_scan(const uint32_t* counts, const uint32_t* order, uint32_t* offsets,
                       uint32_t* ref_counter, uint64_t* status, uint32_t* ticket, uint32_t n,
                       uint32_t epoch, hipStream_t s);
hipError_t launch_emit(const EmitArgs& a, hipStream_t s);
// hist[p][d] += #keys (of n = *n_dev, n <= n_max) with digit d at bit
// shift0 + p*digit_bits. hist_partial: key_hist_rows(n_max) * n_passes * radix words.
uint32_t key_hist_rows(uint32_t n_max);
hipError_t launch_key_histogram(const void* keys, uint32_t key_bytes, const uint32_t* n_dev,
                                uint32_t n_max, uint32_t* hist, uint32_t* hist_partial,
                                uint32_t digit_bits, uint32_t n_passes, uint32_t shift0,
                                hipStream_t s);
// Status words one radix pass needs: radix_status_tiles(n_max) x RADIX.
uint32_t radix_status_tiles(uint32_t n_max);
// One stable LSD pass over n = *n_dev (key, value) pairs, keys of key_bytes
// (4 or 8); with iota_values the input values are the positions 0..n-1 and vin
// is not read. The grid covers n_max; n > n_max is a no-op. With aux_in and
// aux_out (32-bit keys): aux_out[p] = aux_in[vout[p]] (values read from vin),
// or with aux_carry aux_in moves like a second value array: aux_out[p] =
// aux_in[input index of the key that lands at p].
hipError_t launch_radix_pass(const void* kin, const uint32_t* vin, void* kout, uint32_t* vout,
                             uint32_t key_bytes, bool iota_values, const uint32_t* n_dev,
                             uint32_t n_max, const uint32_t* hist_pass, uint64_t* status,
                             uint32_t* ticket, uint32_t shift, uint32_t digit_bits, uint32_t epoch,
                             const uint32_t* aux_in, uint32_t* aux_out, hipStream_t s,
                             const uint32_t* epoch_base = nullptr, uint32_t key_tile = 0,
                             uint32_t match_bits = 0, bool aux_carry = false);
// 64-bit key passes use 16,384-key tiles unless 8,192 is asked for: the
// smaller tiles (two workgroups per CU) measured slower at 3.6M keys (72 vs
// 64 us for two passes) and at 118.9M keys (1.54 vs 1.37 ms).
constexpr uint32_t kSortSmallTileFrom = 0xFFFFFFFFu;
// Digit bits of a pass at `shift` that can be non-zero when keys hold
// `key_bits` significant bits (0 when none).
inline uint32_t pass_match_bits(uint32_t shift, uint32_t digit_bits, uint32_t key_bits) {
    return key_bits <= shift ? 0u : (key_bits - shift < digit_bits ? key_bits - shift : digit_bits);
}
// epoch_base[0] += kEpochStride (the first node of a captured batch graph).
hipError_t launch_bump_epoch(uint32_t* epoch_base, hipStream_t s);
// tile_offsets[0..tile_count] from n = *n_dev sorted keys (n > n_max: all 0):
// 64-bit (tile << 32 | depth) keys, or tile ids alone when key_bytes is 4
// (32-bit) or 2 (16-bit).
// If k_host is set (host-mapped memory), *k_host = *n_dev.
hipError_t launch_identify_tiles(const void* keys, uint32_t key_bytes, const uint32_t* n_dev, uint32_t n_max,
                                 uint32_t* tile_offsets, uint32_t tile_count, uint32_t* k_host,
                                 hipStream_t s);
// keys[i] = (tiles[i] << 32) | bits(-projected[values[i]].depth) for i < *n_dev
// (the 64-bit sort keys, rebuilt from the tile-id sort for readback); tile ids
// of tile_bytes 4 or 2.
hipError_t launch_rebuild_keys(const void* tiles, uint32_t tile_bytes, const uint32_t* values,
                               const gs3d_projected_splat* projected, const uint32_t* n_dev, uint32_t n_max,
                               uint64_t* keys, hipStream_t s);
// Bytes per tile id in the frame's tile-id sort: 2 up to 65,536 tiles, else 4.
inline uint32_t tile_id_bytes(uint32_t tiles) { return tiles <= 65536u ? 2u : 4u; }
// Square tiles of 8, 16 or 32 px; pixels_per_lane 1, 2 or 4 (a workgroup of
// tile*tile/ppl threads, at least one wavefront).
bool raster_variant_exists(uint32_t tile, uint32_t pixels_per_lane);
hipError_t launch_rasterize(const gs3d_projected_splat* projected, const uint32_t* values,
                            const uint32_t* tile_offsets, uint32_t tile_x, uint32_t tile_y,
                            uint32_t tiles_x, uint32_t tiles_y, uint32_t width, uint32_t height,
                            uint32_t pixels_per_lane, uint16_t* image_rgba16f, hipStream_t s,
                            bool quarter_lists = true);
// The same over records of `rec` float4s (4: gs3d_projected_splat; 3: the
// batch's 48-B records, the first three float4s of one). quarter_lists (one
// pixel per lane only): a survivor list per 4 x 4 quarter of each 8 x 8
// block, or one per block.
hipError_t launch_rasterize_records(const float4* records, uint32_t rec, const uint32_t* values,
                                    const uint32_t* tile_offsets, uint32_t tile_x, uint32_t tile_y,
                                    uint32_t tiles_x, uint32_t tiles_y, uint32_t width, uint32_t height,
                                    uint32_t pixels_per_lane, uint16_t* image_rgba16f, hipStream_t s,
                                    bool quarter_lists = true);
hipError_t launch_f16_to_f32(const uint16_t* in, float* out, size_t n, hipStream_t s);

// Launch shape of a pass for a problem size: threads per workgroup and
// workgroups. The launchers size their grids with these same functions.
struct LaunchShape {
    uint32_t block;
    uint64_t grid;
};
LaunchShape preprocess_launch_shape(uint32_t n);
LaunchShape key_histogram_launch_shape(uint32_t n);
LaunchShape radix_pass_launch_shape(uint32_t key_bytes, uint32_t key_tile, uint32_t n);
LaunchShape scan_rects_launch_shape(uint32_t n);
LaunchShape emit_launch_shape(uint32_t n);
LaunchShape identify_launch_shape(uint32_t tiles);
LaunchShape raster_launch_shape(uint32_t tile, uint32_t pixels_per_lane, uint32_t tiles);
LaunchShape composite_launch_shape(uint32_t pixels);

// One view of the batch's persistent rasterizer (device memory).
struct RasterView {
    const float4* proj;            // the view's 48-B records (3 float4s per splat)
    const uint32_t* values;        // its sorted splat ids
    const uint32_t* tile_offsets;  // its tiles + 1 ranges
    uint2* image;                  // its W*H rgba16f target
    uint32_t tiles_x, tiles, width, height;
};
// Control words of the raster queue (zeroed before the build): bucket counts,
// bucket cursors, the rasterizer's ticket.
constexpr uint32_t kRasterQueueCounts = 0;
constexpr uint32_t kRasterQueueCursor = 64;
constexpr uint32_t kRasterQueueTicket = 128;
constexpr uint32_t kRasterQueueWords = 192;
// queue[0 .. sum of tiles) = (view, tile) items, longest tile lists first.
// max_tiles: the largest view's tile count.
hipError_t launch_raster_queue_build(const RasterView* views, uint32_t n_views, uint32_t max_tiles,
                                     uint32_t* ctl, uint2* queue, hipStream_t s);
// Every item of the queue, by a grid of `grid` workgroups (rasterize_queue_grid:
// what the current device holds at once) that take items with a ticket
// (ctl[kRasterQueueTicket]).
// wg_per_cu > 0 caps the workgroups per CU (room for other streams' kernels).
// wave_items: each wavefront takes one pixel block of a tile per ticket
// instead of a workgroup taking the whole tile.
hipError_t rasterize_queue_grid(uint32_t tile, uint32_t pixels_per_lane, uint32_t wg_per_cu, bool wave_items,
                                uint32_t* grid);
hipError_t launch_rasterize_queue(const RasterView* views, const uint2* queue, uint32_t n_items,
                                  uint32_t* ctl, uint32_t tile, uint32_t pixels_per_lane, uint32_t grid,
                                  bool wave_items, hipStream_t s);

// ---- learned colour ---------------------------------------------------------------
struct LearnedColorArgs {
    const float* params;     // GS3D_LEARNED_PARAMS fp32
    const uint16_t* latents; // [n][8] f16
    const float4* pos;
    const float* eyes;       // [n_views][4] camera positions
    uint2* colors;           // [n_views][n] f16 (r, g, b, 1)
    uint32_t n, n_views;
};
hipError_t launch_learned_colors(const LearnedColorArgs& a, hipStream_t s);

struct LearnedTrainArgs {
    const float* params;       // fp32 masters
    const float* latents;      // [n][8] fp32 masters (rounded through f16 on input)
    const float* sh;           // SoA SH coefficients of the splats
    int sh_degree;
    const float* dirs;         // [256][3] unit directions
    const float* basis;        // [256][16] SH basis at each direction (zero-padded)
    uint32_t first_dir;        // this step's directions: first_dir .. first_dir + 15
    uint32_t n;
    float grad_scale;          // d(loss)/d(sum of squared errors)
    float* grad_partial;       // [waves][GS3D_LEARNED_PARAMS]
    float* grad_latents;       // [n][8]
    float* loss_partial;       // [waves] sums of squared errors
};
uint32_t learned_train_waves(uint32_t n);
hipError_t launch_learned_train(const LearnedTrainArgs& a, uint32_t waves, hipStream_t s);
// scratch: learned_reduce_scratch_floats() floats
size_t learned_reduce_scratch_floats();
hipError_t launch_learned_reduce(const float* partial, uint32_t rows, float* grad, const float* loss_partial,
                                 float* loss, float* scratch, hipStream_t s);
hipError_t launch_adam(float* x, float* m, float* v, const float* grad, size_t n, float lr, float b1, float b2,
                       float eps, uint32_t step, uint16_t* f16_copy, hipStream_t s);
// f16_copy[i] = x[i] rounded to binary16 (the latents' f16 copy)
hipError_t launch_f16_copy(const float* x, size_t n, uint16_t* f16_copy, hipStream_t s);
// sum over pixels of the squared RGB differences of two f16 RGBA images (accumulated into *out, f64)
hipError_t launch_sq_diff_sum(const uint16_t* a, const uint16_t* b, size_t pixels, double* out, hipStream_t s);

// Composite of the f16 target over an unorm8 framebuffer, through the
// splats-per-tile heat map when debug is set (then also written to
// debug_target).
struct CompositeParams {
    const uint16_t* target;
    uint16_t* debug_target;
    uint32_t* framebuffer;
    const uint32_t* tile_offsets;
    uint32_t width, height, tile_x, tile_y, tiles_x, tiles_y;
    uint32_t debug_ref;
    uint32_t clear;  // packed unorm8 clear colour (framebuffer byte order)
    int use_clear;
    int bgra;
    int debug;
};
// Packs an RGBA float colour into one unorm8 framebuffer word (the clear pass).
uint32_t pack_unorm8_clear(const float rgba[4], int bgra);
hipError_t launch_composite(const CompositeParams& p, hipStream_t s);

}  // namespace gs3d
