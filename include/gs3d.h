/*
 * gs3d.h — C ABI of libgs3d.so, the MI355X (gfx950) rebin-then-rasterize
 * 3D Gaussian Splatting path.
 *
 * Every entry point returns GS3D_OK (0) or a negative GS3D_E* code; the message
 * of the last failure on a context is gs3d_last_error(ctx). No C++ exception
 * crosses this boundary. Host pointers are borrowed for the duration of the
 * call only (copied host->device or device->host inside the call). A context
 * owns all of its device memory and one HIP stream; it is not thread-safe
 * (one caller thread per context, as the reference's single JS thread).
 *
 * Reference interfaces replaced (paths relative to the reference repo):
 *   IRenderer.init / render ............ src/renderers/IRenderer.ts:10-13
 *   GaussianSplatRenderer.rebin ........ src/renderers/GaussianSplatRenderer.ts:1104-1266
 *   setTileSize / getTileCount / Size .. GaussianSplatRenderer.ts:301-317, 1268-1274
 *   readTileCounts ..................... GaussianSplatRenderer.ts:1278-1310
 *   SceneSyncer.setSplatData ........... src/SceneSyncer.ts:25-43
 *   BufferManager named buffers ........ src/BufferManager.ts:32-154
 *   Profiler per-pass timings .......... src/Profiler.ts:135-251
 *   WGSL passes (stage entry points) ... src/shaders/compute/ (one per pass)
 */
#ifndef GS3D_H
#define GS3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS3D_ABI_VERSION 1

/* ---- error codes -------------------------------------------------------- */
#define GS3D_OK               0
#define GS3D_E_INVALID_ARG   -1   /* null pointer, zero/oversize dimension      */
#define GS3D_E_NO_DEVICE     -2   /* HIP device missing or not gfx950-capable   */
#define GS3D_E_HIP           -3   /* a HIP runtime call failed                  */
#define GS3D_E_OOM           -4   /* device allocation failed                   */
#define GS3D_E_STATE         -5   /* call order: e.g. rebin before set_splats   */
#define GS3D_E_UNKNOWN_NAME  -6   /* registry miss (buffer / kernel variant)    */
#define GS3D_E_TOO_SMALL     -7   /* destination buffer smaller than the data   */
#define GS3D_E_FORMAT        -8   /* malformed or unsupported PLY file          */
#define GS3D_E_IO            -9   /* a file could not be opened or read         */

/* ---- byte-exact data contract ------------------------------------------- */

/* CameraUniforms, 72 x f32 = 288 B: view, projection, inverse view, inverse
 * projection (column-major mat4), position.xyz1, viewport (w, h, w/h, 0).
 * Reference: preprocess_splats.wgsl:1-8, Camera.getUniformData Camera.ts:172-191. */
typedef struct gs3d_camera {
    float view[16];
    float projection[16];
    float inverse_view[16];
    float inverse_projection[16];
    float position[4];
    float viewport[4];
} gs3d_camera;

/* ProjectedSplat, 64 B. Reference: preprocess_splats.wgsl:18-31.
 * depth is the view-space z (negative in front of the camera). */
typedef struct gs3d_projected_splat {
    float    mean_px[2];
    float    depth;
    float    radius_px;
    float    conic[3];
    float    opacity;
    float    color[3];
    uint32_t valid;
    uint32_t tile_min[2];
    uint32_t tile_max[2];
} gs3d_projected_splat;

/* A sort key is the little-endian u64 (tile_id << 32) | bits(-depth), i.e. the
 * interleaved {hi, lo} u32 pair of emit_tile_refs.wgsl:41-46 read as one u64. */

/* Buffers readable with gs3d_read (the reference's named buffers,
 * GaussianSplatRenderer.ts:338-465). */
enum gs3d_buffer_id {
    GS3D_BUF_PROJECTED_SPLATS  = 0,  /* N x gs3d_projected_splat                    */
    GS3D_BUF_SPLAT_REF_COUNTS  = 1,  /* N x u32                                     */
    GS3D_BUF_SPLAT_REF_OFFSETS = 2,  /* (N+1) x u32, [N] = total (sentinel)         */
    GS3D_BUF_REF_COUNTER       = 3,  /* 1 x u32 = K                                 */
    GS3D_BUF_SORT_KEYS         = 4,  /* K x u64, sorted by (tile, depth)            */
    GS3D_BUF_SORT_VALUES       = 5,  /* K x u32 splat ids, in sorted order          */
    GS3D_BUF_TILE_OFFSETS      = 6,  /* (T+1) x u32, trailing tiles stay 0          */
    GS3D_BUF_IMAGE_RGBA16F     = 7,  /* W*H x 4 x f16, premultiplied (C, 1-T)       */
    GS3D_BUF_IMAGE_RGBA32F     = 8,  /* same image widened to f32 on readback       */
    GS3D_BUF_UNSORTED_KEYS     = 9,  /* K x u64 emit order (before the sort)        */
    GS3D_BUF_UNSORTED_VALUES   = 10, /* K x u32 emit order                          */
    GS3D_BUF_DEBUG_TARGET_RGBA16F = 11, /* W*H x 4 x f16: target + heat map (last
                                           composite in debug mode 1)              */
    GS3D_BUF_FRAMEBUFFER       = 12, /* W*H x 4 unorm8 of the last gs3d_composite  */
    GS3D_BUF__COUNT            = 13
};

/* Framebuffer formats of gs3d_composite (the swapchain format the reference
 * gets from navigator.gpu.getPreferredCanvasFormat, DeviceManager.ts:15). */
enum gs3d_format {
    GS3D_FORMAT_RGBA8UNORM = 0,
    GS3D_FORMAT_BGRA8UNORM = 1
};

/* rebin flags */
#define GS3D_REBIN_PROFILE     1u   /* record a hipEvent pair per stage           */
#define GS3D_REBIN_KEEP_UNSORTED 2u /* keep a copy of the emit-order keys/values  */

/* Stage labels of gs3d_stage_times_us, in execution order (reference Profiler
 * labels where a reference pass exists, GaussianSplatRenderer.ts:1143-1265). */
enum gs3d_stage {
    GS3D_STAGE_PREPROCESS = 0,  /* "preprocess-splats"                            */
    GS3D_STAGE_DEPTH_SORT = 1,  /* "radix-sort-splats": depth digits, on splats   */
    GS3D_STAGE_SCAN       = 2,  /* "prefix-scan"  (local+blocks+add fused)        */
    GS3D_STAGE_EMIT       = 3,  /* "emit-tile-refs" (+ tile-digit histograms)     */
    GS3D_STAGE_SORT       = 4,  /* "radix-sort": tile digits, on the keys         */
    GS3D_STAGE_IDENTIFY   = 5,  /* "identify-tile-ranges"                         */
    GS3D_STAGE_RASTER     = 6,  /* "rasterize-tiles"                              */
    GS3D_STAGE_TOTAL      = 7,  /* whole rebin, first launch to last              */
    GS3D_STAGE_COMPOSITE  = 8,  /* "composite-splats": the display pass of
                                   gs3d_composite ("tile-overlay" fused in, in
                                   debug mode 1), timed after a profiled rebin   */
    GS3D_STAGE__COUNT     = 9
};

typedef struct gs3d_ctx gs3d_ctx;

/* ---- context ------------------------------------------------------------ */

/* Create a context on HIP device `device` (IRenderer.init, IRenderer.ts:10-11). */
int gs3d_create(int device, gs3d_ctx** out);
int gs3d_destroy(gs3d_ctx* ctx);
const char* gs3d_last_error(const gs3d_ctx* ctx);
int gs3d_abi_version(void);
/* Block until all work queued on the context's stream has finished. */
int gs3d_sync(gs3d_ctx* ctx);

/* Device limits in the reference's WorkgroupLimits order
 * (types.ts:49-56, WorkgroupManager.ts:15-24): maxTotalThreads, maxSizeX,
 * maxSizeY, maxSizeZ, maxDispatch, maxSharedMemory. */
int gs3d_device_limits(gs3d_ctx* ctx, uint32_t out[6]);

/* ---- scene upload (SceneSyncer.setSplatData, SceneSyncer.ts:25-43) ---------
 * pos4/scale4/rot4/color4: N x vec4 f32 in the loader's layouts
 * (SceneWorker.ts:407-464): (x,y,z,1), (e^s0,e^s1,e^s2,0), normalised (w,x,y,z),
 * (clamp01(0.5+C0 f_dc), clamp01(sigmoid(opacity))).
 * sh_degree 0 uses color4 directly (the reference behaviour). For sh_degree
 * 1..3, f_dc (N x 3 raw PLY values) and sh_rest (N x 3*((d+1)^2-1) f32 in
 * f_rest_k index order: all R coefficients, then G, then B) are required. */
int gs3d_set_splats(gs3d_ctx* ctx,
                    const float* pos4, const float* scale4,
                    const float* rot4, const float* color4,
                    const float* f_dc, const float* sh_rest,
                    int sh_degree, uint32_t n);
uint32_t gs3d_splat_count(const gs3d_ctx* ctx);

/* ---- renderer API (GaussianSplatRenderer.ts) ----------------------------- */

/* Clamp to device limits and set the tile size; writes the size in use
 * (setTileSize, GaussianSplatRenderer.ts:301-317, clampTileSize :162-167).
 * Compiled tile variants are 8x8, 16x16 and 32x32 (1 thread per pixel up to
 * 1024 threads); other requests are rejected with GS3D_E_UNKNOWN_NAME. */
int gs3d_set_tile_size(gs3d_ctx* ctx, uint32_t x, uint32_t y, uint32_t out_xy[2]);
int gs3d_get_tile_size(const gs3d_ctx* ctx, uint32_t out_xy[2]);

/* Compiled kernel variants (PipelineManager.create/update with codeConstants,
 * src/PipelineManager.ts:34-105; the WGSL is specialised by text substitution,
 * here a precompiled variant is selected):
 *   "rasterize-tiles" "PIXELS_PER_LANE"  1, 2, 4 (workgroup tile^2/ppl >= 64)
 *   "rasterize-tiles" "QUARTER_LISTS"    2 (default), 1, 0: at one pixel per
 *                                        lane, a survivor list per 4x4 quarter
 *                                        of each 8x8 pixel block (1), or one per
 *                                        block (0, faster when most splats span
 *                                        the whole block), or (2) per frame:
 *                                        one per block when the previous frame
 *                                        had more than 8 references per splat
 *                                        at 16x16 tiles (scaled by tile area),
 *                                        else per quarter (batches: per
 *                                        quarter); the image is the same
 *   "radix-sort"      "RADIX_BITS"       4, 8
 *   "radix-sort"      "KEY_TILE"         0 (16384), 8192, 16384: keys per
 *                                        workgroup of the tile-id passes (and of
 *                                        gs3d_stage_sort's 64-bit passes)
 *   "radix-sort-splats" "KEY_TILE"       0 (8192), 4096: keys per workgroup of
 *                                        the 32-bit depth passes
 *   "rasterize-tiles-batch" (gs3d_rebin_batch):
 *                     "GROUPS"           1 (default): one persistent rasterizer
 *                                        over every view's tiles after the
 *                                        views' binning; g > 1: g such launches
 *                                        over view groups; 0: a raster launch per
 *                                        view on its binning stream
 *                     "WAVE_ITEMS"       0 (default), 1: the persistent
 *                                        rasterizer takes pixel blocks instead
 *                                        of tiles per ticket
 *                     "WG_PER_CU"        0 (as many as fit) .. 16: its
 *                                        workgroups per CU
 * get_kernel_variant also reads "rasterize-tiles" "TILE_X". Anything else:
 * GS3D_E_UNKNOWN_NAME. A tile-size change resets an impossible PIXELS_PER_LANE
 * to 1. */
int gs3d_set_kernel_variant(gs3d_ctx* ctx, const char* kernel, const char* param, int value);
int gs3d_get_kernel_variant(gs3d_ctx* ctx, const char* kernel, const char* param, int* value);
int gs3d_get_tile_count(const gs3d_ctx* ctx, uint32_t out_xy[2]);

/* Stages 1-6A on the context's stream: preprocess, scan, emit, sort, identify,
 * rasterize (rebin, GaussianSplatRenderer.ts:1104-1266). `camera` is the 288-B
 * CameraUniforms. Returns once the work is queued; gs3d_read / gs3d_sync wait. */
int gs3d_rebin(gs3d_ctx* ctx, const gs3d_camera* camera,
               uint32_t width, uint32_t height, uint32_t flags);

/* Number of keys K emitted by the last rebin (ref_counter[0]). */
int gs3d_ref_count(gs3d_ctx* ctx, uint64_t* out);

/* Copy a named result of the last rebin into caller memory. `written` gets the
 * number of bytes the buffer holds; GS3D_E_TOO_SMALL if `bytes` is short. */
int gs3d_read(gs3d_ctx* ctx, int buffer_id, void* dst, size_t bytes, size_t* written);

/* Per-tile splat counts with the underflow guard of readTileCounts
 * (GaussianSplatRenderer.ts:1278-1310): counts[t] = off[t+1] >= off[t] ?
 * off[t+1]-off[t] : 0 for t < T. */
int gs3d_read_tile_counts(gs3d_ctx* ctx, uint32_t* counts, size_t n, size_t* written);

/* Per-stage GPU times (microseconds) of the last rebin with GS3D_REBIN_PROFILE,
 * indexed by enum gs3d_stage (Profiler.ts:197-228 sums repeated labels). */
int gs3d_stage_times_us(gs3d_ctx* ctx, float* out, int n);
const char* gs3d_stage_label(int stage);

/* Which stages do the work of one pass of the reference's Profiler table,
 * by its label (GaussianSplatRenderer.ts:1017-1260): a bit mask, bit i = enum
 * gs3d_stage i; 0 for a label the reference does not time. Several reference
 * passes map to one fused stage here ("prefix-scan-local/-blocks/-add" ->
 * prefix-scan; the histogram, table scans and scatter of the radix passes ->
 * radix-sort-splats for the 32 depth bits and radix-sort for the tile bits;
 * "tile-overlay" and "composite-splats" -> composite-splats). */
uint32_t gs3d_reference_pass_stages(const char* reference_label);

/* Average per-stage times over every profiled rebin since the last reset
 * (the Profiler's averaging window, Profiler.ts:231-251); `frames` gets the
 * number of frames averaged. */
int gs3d_profiler_reset(gs3d_ctx* ctx);
/* Which stages a GS3D_REBIN_PROFILE rebin brackets with events (bit i = enum
 * gs3d_stage i; default all). Each event pair adds a dispatch gap, so a caller
 * timing whole frames profiles only what it needs. Unrecorded stages read 0. */
int gs3d_set_profile_mask(gs3d_ctx* ctx, uint32_t stage_mask);
int gs3d_stage_times_avg_us(gs3d_ctx* ctx, float* out, int n, uint64_t* frames);

/* Launch geometry (WorkgroupManager.getLayout, src/WorkgroupManager.ts:58-147:
 * the reference computes workgroup and dispatch sizes from device limits;
 * here the library owns its launches and reports them). For one launch of
 * `kernel` over a problem of problem[0] x problem[1] x problem[2]: threads per
 * workgroup in workgroup[0..2] and workgroups in grid[0..2], exactly as the
 * library launches it. Kernels and their problem sizes:
 *   "preprocess-splats", "radix-histogram" (the depth-key digit histogram),
 *   "radix-sort-splats" (one depth pass), "prefix-scan", "emit-tile-refs":
 *       problem[0] = splats
 *   "radix-sort" (one tile-id pass): problem[0] = keys
 *   "identify-tile-ranges", "rasterize-tiles", "composite-splats":
 *       problem[0], problem[1] = viewport width, height
 * ctx may be NULL (the defaults: 16x16 tiles, 1 pixel per lane, default key
 * tiles); otherwise the context's tile size and kernel variants count.
 * GS3D_E_UNKNOWN_NAME for another kernel name. */
int gs3d_launch_geometry(const gs3d_ctx* ctx, const char* kernel, const uint64_t problem[3],
                         uint32_t workgroup[3], uint64_t grid[3]);

/* Number of radix passes the last rebin's sort ran and the bit offset of each
 * (passes whose digit is constant over all keys are skipped). */
int gs3d_sort_passes(gs3d_ctx* ctx, uint32_t* n_passes, uint32_t bit_offsets[16]);

/* ---- named device buffers (BufferManager, src/BufferManager.ts:32-154) ----
 * The renderer's own buffers ("projected_splats", "splat_ref_counts",
 * "splat_ref_offsets", "sort_keys_a", "sort_values_a", "tile_offsets",
 * "splat_target", ...) live in the same registry. */
/* initBuffers/createBuffer: allocate (at least 4 bytes); an existing name is
 * resized. Contents are undefined after a (re)allocation. */
int gs3d_buffer_create(gs3d_ctx* ctx, const char* name, size_t bytes);
/* resize: GS3D_E_UNKNOWN_NAME for a name never created. */
int gs3d_buffer_resize(gs3d_ctx* ctx, const char* name, size_t bytes);
int gs3d_buffer_size(gs3d_ctx* ctx, const char* name, size_t* bytes);
/* write: clipped to the buffer (BufferManager.ts:92-125); when nothing fits it
 * is skipped with GS3D_E_TOO_SMALL. `written` gets the bytes copied. */
int gs3d_buffer_write(gs3d_ctx* ctx, const char* name, size_t offset, const void* src,
                      size_t bytes, size_t* written);
int gs3d_buffer_read(gs3d_ctx* ctx, const char* name, size_t offset, void* dst, size_t bytes);
/* destroy: frees the buffer; an unknown name is a no-op (BufferManager.ts:164-170). */
int gs3d_buffer_destroy(gs3d_ctx* ctx, const char* name);
/* All buffers, largest first: names[i*64 .. i*64+63], sizes[i], for i < n_max;
 * `count` gets the total number (Profiler.getBuffersSortedBySize). */
int gs3d_buffer_list(gs3d_ctx* ctx, char* names, uint64_t* sizes, size_t n_max, size_t* count);

/* ---- batched views ------------------------------------------------------------
 * Many views of the resident splats in one call (no reference counterpart:
 * the reference renders one view per frame). Each view gets the result a
 * gs3d_rebin of that camera and viewport would give. Views may differ in
 * resolution; the tile size, digit width and raster variant are the
 * context's. The launches are captured into one hipGraph on first use and
 * replayed while the view shapes stay the same (cameras may change freely);
 * GS3D_BATCH_EAGER launches them directly instead. The call returns when
 * every view is done. */
#define GS3D_MAX_BATCH_VIEWS 1024u
#define GS3D_BATCH_EAGER 1u

typedef struct gs3d_view {
    gs3d_camera camera;      /* 288-B camera uniform (CameraUniforms)      */
    uint32_t    width, height;
} gs3d_view;

int gs3d_rebin_batch(gs3d_ctx* ctx, const gs3d_view* views, uint32_t n_views, uint32_t flags);
/* K (keys) of every view of the last batch. */
int gs3d_batch_ref_counts(gs3d_ctx* ctx, uint64_t* out, uint32_t n);
/* Per-view results of the last batch: GS3D_BUF_IMAGE_RGBA16F / _RGBA32F,
 * GS3D_BUF_TILE_OFFSETS, GS3D_BUF_PROJECTED_SPLATS, GS3D_BUF_REF_COUNTER,
 * GS3D_BUF_SORT_VALUES and GS3D_BUF_SORT_KEYS (the keys rebuilt from the sorted
 * ids, the tile ranges and the view's projected depths).
 * A batch keeps only what its rasterizer reads of the projected splats; the
 * N x gs3d_projected_splat of a view are projected again on read, so they
 * need the batch's splats: GS3D_E_STATE after a new gs3d_set_splats. */
int gs3d_batch_read(gs3d_ctx* ctx, uint32_t view, int id, void* dst, size_t bytes, size_t* written);

/* ---- learned colour -----------------------------------------------------------
 * No reference counterpart. Each splat stores an 8-dim fp16 latent z; a shared
 * MLP maps (z, unit view direction) -> 64 -> 64 -> 3 (ReLU, ReLU, sigmoid)
 * with bf16 weights and fp32 accumulation on MFMA. It is trained on the GPU
 * against the exact SH colours clamp01(0.5 + sum_k b_k(d) c_k) of the uploaded
 * splats at 256 fixed unit directions (a Fibonacci sphere) with Adam, and
 * replaces the SH evaluation in a batch rendered with GS3D_BATCH_LEARNED once
 * gs3d_learned_validate has found it within the PSNR threshold of exact SH;
 * otherwise such a batch falls back to exact SH.
 * Parameters, packed fp32: W1[64][11], b1[64], W2[64][64], b2[64], W3[3][64],
 * b3[3] (the input vector is z0..z7, dx, dy, dz). */
#define GS3D_LEARNED_PARAMS 5123
#define GS3D_LEARNED_DIRS 256
#define GS3D_BATCH_LEARNED 2u

/* Initial parameters (He-normal weights, zero biases, latents N(0, 0.1)) from `seed`. */
int gs3d_learned_init(gs3d_ctx* ctx, uint64_t seed);
/* `epochs` passes over splats x 256 directions in steps of 16 directions;
 * Adam (b1 0.9, b2 0.999, eps 1e-8) with learning rates lr (weights) and
 * lr_latent (latents). loss[e] (nullable, n_loss entries) = mean squared error
 * of epoch e. */
int gs3d_learned_train(gs3d_ctx* ctx, uint32_t epochs, float lr, float lr_latent,
                       float* loss, uint32_t n_loss);
/* Gradients of the mean squared error over splats x directions
 * [first_dir, first_dir + 16) at the current parameters, without an update. */
int gs3d_learned_grad(gs3d_ctx* ctx, uint32_t first_dir, float* grad_params,
                      float* grad_latents, float* loss);
int gs3d_learned_get(gs3d_ctx* ctx, float* params, float* latents, uint16_t* latents_f16);
int gs3d_learned_set(gs3d_ctx* ctx, const float* params, const float* latents);
/* The whole training state, to checkpoint and resume training: the
 * parameters and latents (layouts of gs3d_learned_get), the Adam first and
 * second moments of each, and the Adam step count (bias correction). Any
 * array pointer may be NULL to skip that array. Setting a state needs
 * gs3d_learned_init on the same splats first; it re-derives the f16 latent
 * copy and leaves learned colours disabled until gs3d_learned_validate. */
typedef struct gs3d_learned_state {
    float* params;        /* GS3D_LEARNED_PARAMS */
    float* params_m;
    float* params_v;
    float* latents;       /* N x 8 */
    float* latents_m;
    float* latents_v;
    uint32_t step;        /* Adam steps taken */
} gs3d_learned_state;
int gs3d_learned_get_state(gs3d_ctx* ctx, gs3d_learned_state* state);
int gs3d_learned_set_state(gs3d_ctx* ctx, const gs3d_learned_state* state);
/* The 256 training directions ([256][3]). */
int gs3d_learned_dirs(float* dirs);
/* Learned colours of every splat seen from `eye` (N x 3 f32). */
int gs3d_learned_colors(gs3d_ctx* ctx, const float eye[3], float* rgb);
/* Render the views with exact SH and with learned colours; psnr = 10 log10(1 /
 * MSE) over the RGB of every pixel of every view. Learned colours are enabled
 * for GS3D_BATCH_LEARNED iff psnr >= min_psnr_db. */
int gs3d_learned_validate(gs3d_ctx* ctx, const gs3d_view* views, uint32_t n_views,
                          float min_psnr_db, float* psnr, int* enabled);
/* Force the learned colours on or off for GS3D_BATCH_LEARNED batches. */
int gs3d_learned_enable(gs3d_ctx* ctx, int on);

/* ---- display: debug heat map and composite ---------------------------------- */

/* setDebugMode (GaussianSplatRenderer.ts:319-323): 0 = normal, 1 = the
 * splats-per-tile heat map (debug_tile_overlay.wgsl:35-58) over the target.
 * debug_ref is the splat count drawn full red (the reference's debugRef, 256,
 * :75). Any mode other than 0 or 1 is GS3D_E_INVALID_ARG. */
int gs3d_set_debug_mode(gs3d_ctx* ctx, uint32_t mode, uint32_t debug_ref);
int gs3d_get_debug_mode(const gs3d_ctx* ctx, uint32_t* mode, uint32_t* debug_ref);

/* The composite pass of render() (GaussianSplatRenderer.ts:1075-1097): the
 * last rebin's premultiplied target, through the heat map in debug mode 1,
 * blended "over" (one, one-minus-src-alpha on colour and alpha, :939-945)
 * onto a W*H unorm8 framebuffer in `format` byte order (composite_splats.wgsl:12-29).
 * The framebuffer is the context's resident buffer "framebuffer":
 *   clear_rgba non-null: it is first cleared to that colour (the viewport's
 *     clear pass, Viewport.ts:146-154; the reference clears to 0.12, 0.12, 0.13, 1);
 *   else, framebuffer non-null: the caller's W*H*4 bytes are uploaded first;
 *   else the resident contents are blended onto (undefined after the
 *     viewport size changed).
 * With `framebuffer` non-null the result is copied back into it (`bytes` must
 * hold W*H*4). Unorm stores round to nearest: floor(clamp(v, 0, 1)*255 + 0.5). */
int gs3d_composite(gs3d_ctx* ctx, int format, const float* clear_rgba,
                   uint8_t* framebuffer, size_t bytes);

/* ---- native PLY ingest (SceneWorker.ts:146-549, SceneSyncer.ts:25-43) ---------
 * The reference parses PLY files in a Web Worker in JS (parsePlyHeader,
 * createPlyValueReader, parseGaussianSplatBody, writeGaussianSplat). These
 * entry points restate it in C++: same header rules and error messages, same
 * activations (f64, stored as f32), binary bodies parsed by several threads.
 * Divergences: element counts must be whole numbers below 2^32; ascii tokens
 * are split on ASCII whitespace only. */
enum gs3d_ply_format {
    GS3D_PLY_ASCII     = 0,
    GS3D_PLY_BINARY_LE = 1,
    GS3D_PLY_BINARY_BE = 2
};

typedef struct gs3d_ply_info {
    uint64_t splat_count;    /* rows of the vertex element                         */
    uint32_t f_rest_count;   /* f_rest_k properties (sorted by k)                  */
    int32_t  sh_degree;      /* inferSphericalHarmonicsDegree(f_rest_count), :501  */
    int32_t  format;         /* gs3d_ply_format                                    */
    uint32_t reserved;
    uint64_t header_bytes;   /* body offset                                        */
    double   bounds_min[3];  /* min / max of x, y, z (after gs3d_ply_parse)        */
    double   bounds_max[3];
} gs3d_ply_info;

/* Header only: fills splat_count, f_rest_count, sh_degree, format, header_bytes.
 * GS3D_E_FORMAT with the reference's message in `err` on a bad header. */
int gs3d_ply_probe(const void* data, size_t bytes, gs3d_ply_info* info,
                   char* err, size_t err_bytes);

/* Whole file into the loader's arrays (GaussianSplatData, SceneWorker.ts:389-405):
 * positions/scales/rotations/colors N x 4 f32, f_dc N x 3 f32 (raw values), and
 * sh_rest N x f_rest_count f32 (null allowed when f_rest_count is 0).
 * threads: 0 = up to 16. No device or context is needed. */
int gs3d_ply_parse(const void* data, size_t bytes, float* positions, float* scales,
                   float* rotations, float* colors, float* f_dc, float* sh_rest,
                   int threads, gs3d_ply_info* info, char* err, size_t err_bytes);

/* Load a PLY file and upload it (gs3d_set_splats) through pinned staging
 * memory. SH degree d uses the first (d+1)^2-1 coefficients of each colour
 * channel when the file holds more. */
int gs3d_load_ply(gs3d_ctx* ctx, const char* path, int threads, gs3d_ply_info* info);

/* ---- stage entry points (each WGSL pass fed caller data) ---------------- */

/* preprocess_splats.wgsl:75-184 over caller splats. projected: N x 64 B,
 * counts: N x u32. tile_count/inv_tile are SplatBinningUniforms fields
 * (GaussianSplatRenderer.ts:1119-1131). Uses the uploaded SH if sh_degree>0. */
int gs3d_stage_preprocess(gs3d_ctx* ctx, const gs3d_camera* camera,
                          uint32_t tiles_x, uint32_t tiles_y,
                          float inv_tile_x, float inv_tile_y,
                          gs3d_projected_splat* projected, uint32_t* counts);

/* Exclusive prefix scan of u32 counts (prefix_scan_local/blocks/add.wgsl):
 * offsets[i] = sum_{j<i} counts[j] mod 2^32, offsets[n] = total. `offsets`
 * holds n+1 entries. */
int gs3d_stage_scan(gs3d_ctx* ctx, const uint32_t* counts, uint32_t n,
                    uint32_t* offsets, uint32_t* total);

/* emit_tile_refs.wgsl:56-79: one (key, splat_id) per overlapped tile, written
 * at offsets[id]. `offsets` has n+1 entries; keys/values hold offsets[n]. */
int gs3d_stage_emit(gs3d_ctx* ctx, const gs3d_projected_splat* projected,
                    const uint32_t* offsets, uint32_t n, uint32_t tiles_x,
                    uint64_t* keys, uint32_t* values);

/* Stable LSD radix sort of (u64 key, u32 value) pairs — the result of
 * radix_histogram/scan_local/scan_add/radix_scatter.wgsl over all 64 bits:
 * pairs ordered by (key, input position). digit_bits is 4 or 8. */
int gs3d_stage_sort(gs3d_ctx* ctx, const uint64_t* keys_in, const uint32_t* values_in,
                    uint32_t n, uint32_t digit_bits,
                    uint64_t* keys_out, uint32_t* values_out);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the sort of gs3d_stage_sort (same order, same digit
 * widths) over caller DEVICE buffers, queued on the context's stream, with no
 * allocation, free, copy or synchronisation. Everything beyond the inputs and
 * outputs (histograms, lookback words, the second ping-pong buffer) lives in
 * a static device array of about 58 MiB that each GPU the library runs on
 * holds until the process ends. n above GS3D_SORT_NOALLOC_MAX_PAIRS is
 * GS3D_E_INVALID_ARG. keys_out/values_out may be keys_in/values_in. The
 * results are ready after gs3d_sync or any later blocking call on the context.
 * Every call on one GPU shares that GPU's array, from every context: calls
 * that overlap on one GPU (two contexts' streams, or two threads) race. */
#define GS3D_SORT_NOALLOC_MAX_PAIRS (1u << 22)
int gs3d_sort_pairs_noalloc(gs3d_ctx* ctx, const uint64_t* keys_in, const uint32_t* values_in,
                            uint32_t n, uint32_t digit_bits, uint64_t* keys_out, uint32_t* values_out);

/* identify_tile_ranges.wgsl:30-52 over caller sorted keys; tile_offsets holds
 * tile_count+1 entries, zero-initialised inside the call. */
int gs3d_stage_identify(gs3d_ctx* ctx, const uint64_t* sorted_keys, uint32_t n,
                        uint32_t tile_count, uint32_t* tile_offsets);

/* rasterize_tiles.wgsl:43-110 with tile size tile_x x tile_y (8, 16 or 32).
 * image: W*H*4 f32 premultiplied (C, 1-T), rounded through f16 like the
 * reference's rgba16float target. */
int gs3d_stage_rasterize(gs3d_ctx* ctx, const gs3d_projected_splat* projected,
                         uint32_t n_splats, const uint32_t* sorted_values, uint32_t n_refs,
                         const uint32_t* tile_offsets, uint32_t tile_x, uint32_t tile_y,
                         uint32_t width, uint32_t height, float* image);

/* The heat map and composite over caller data: target_f16 W*H*4 binary16
 * halves (premultiplied), tile_offsets (T+1) u32 for tile_x x tile_y tiles
 * (read only when debug is 1), framebuffer W*H*4 unorm8 read and written (its
 * input ignored when clear_rgba is set). debug_out (nullable) receives the
 * W*H*4 f16 heat-map target when debug is 1. */
int gs3d_stage_composite(gs3d_ctx* ctx, const uint16_t* target_f16,
                         const uint32_t* tile_offsets, uint32_t tile_x, uint32_t tile_y,
                         uint32_t width, uint32_t height, uint32_t debug, uint32_t debug_ref,
                         int format, const float* clear_rgba, uint8_t* framebuffer,
                         uint16_t* debug_out);

#ifdef __cplusplus
}
#endif
#endif /* GS3D_H */
