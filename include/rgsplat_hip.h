/* rgsplat_hip.h -- C ABI of the MI355X (gfx950) Gaussian-splatting training
 * step. Drop-in for the Rayon CPU path of the rgsplat crate: the record
 * layouts are byte-compatible with its #[repr(C)] types, and each entry point
 * below names the crate function it replaces.
 *
 * Conventions: every function returns 0 on success and a nonzero RGS_E_* code
 * on failure (never aborts, never unwinds); rgs_last_error() describes the
 * last failure on the calling thread. A context is single-threaded: it owns
 * one HIP stream and its device buffers. Host pointers are borrowed for the
 * duration of the call; results are complete when the call returns.
 */
#ifndef RGSPLAT_HIP_H
#define RGSPLAT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RGS_OK = 0,
  RGS_E_ARG = 1,      /* invalid argument (null pointer, size mismatch) */
  RGS_E_HIP = 2,      /* HIP runtime failure */
  RGS_E_CAMERA = 3,   /* c2w not invertible (math/mod.rs:155-160 panics) */
  RGS_E_NAN = 4,      /* NaN depth (backward.rs:191 panics in the sort) */
  RGS_E_CAPACITY = 5, /* densify result exceeds the caller's capacity, or a
                         render needs 2^31 - 1 or more (tile, splat) pairs */
  RGS_E_IO = 6,       /* a model file could not be opened, read or written */
  RGS_E_FORMAT = 7,   /* not a PLY the reference reads (io/ply.rs:116-170 bails) */
  RGS_E_HOST = 8      /* a host-side failure, e.g. host memory exhausted */
};

/* Gaussian3D (math/mod.rs:27-46), 59 x f32 = 236 bytes. */
typedef struct rgs_gaussian {
  float position[3];
  float opacity_logit;
  float rotation[4]; /* w, x, y, z */
  float log_scale[3];
  float sh_coeffs[48]; /* channel-major: sh[ch * 16 + k] */
} rgs_gaussian;

/* GaussianGrad (backward.rs:33-41), 60 x f32 = 240 bytes. Field order equals
 * the optimizer's flat parameter order plus screen_grad_norm. */
typedef struct rgs_grad {
  float d_position[3];
  float d_rotation[4];
  float d_log_scale[3];
  float d_opacity_logit;
  float d_sh[48];
  float screen_grad_norm;
} rgs_grad;

/* Camera (math/mod.rs:130-145); c2w row-major, OpenCV convention. */
typedef struct rgs_camera {
  uint32_t width, height;
  float fx, fy, cx, cy;
  float c2w[16];
  float distortion[3]; /* ignored, as in the reference */
} rgs_camera;

/* TrainConfig (pipeline/optimizer.rs:23-40). */
typedef struct rgs_train_cfg {
  float lr_position, lr_rotation, lr_scale, lr_opacity, lr_sh_dc, lr_sh_rest;
  float beta1, beta2, eps;
  float densify_grad_threshold, prune_opacity_thresh;
  uint32_t densify_from_iter, densify_until_iter, opacity_reset_iter;
  uint64_t max_gaussians, num_iterations;
} rgs_train_cfg;

/* AdamState (pipeline/optimizer.rs:74-84) in host memory: m1/m2 are
 * rows x 59, grad_accum/grad_denom have `rows` entries; buffers have room
 * for `capacity` rows. rows only grows (AdamState::resize, :97-105). */
typedef struct rgs_adam {
  float* m1;
  float* m2;
  float* grad_accum;
  uint32_t* grad_denom;
  size_t rows;
  size_t capacity;
  uint64_t t;
} rgs_adam;

typedef struct rgs_ctx rgs_ctx;

/* Context lifecycle (no reference counterpart: the crate uses the global rayon
 * pool, pipeline/mod.rs:191-194). device < 0 selects the current device. */
int rgs_create(int device, rgs_ctx** out);
void rgs_destroy(rgs_ctx* ctx);
const char* rgs_last_error(void);
/* Run all later work on an external HIP stream, e.g. a framework's current
 * stream; NULL is the HIP null stream (a framework's default stream, whose
 * handle is 0). rgs_train_step does not capture graphs on the null stream.
 * rgs_use_own_stream returns to the context's own (non-blocking) stream. */
int rgs_set_stream(rgs_ctx* ctx, void* hip_stream);
int rgs_use_own_stream(rgs_ctx* ctx);

/* render_and_backward (cpu/backward.rs:107-149). target_rgba is w*h*4 f32;
 * grads_out has n entries, zero for culled Gaussians. */
int rgs_render_and_backward(rgs_ctx* ctx, const rgs_gaussian* g, size_t n, const rgs_camera* cam,
                            const float* target_rgba, const float bg[3], float lambda_dssim, float* loss_out,
                            rgs_grad* grads_out);

/* render_color (cpu/backward.rs:152-166): rgb_out is h*w*3. */
int rgs_render_color(rgs_ctx* ctx, const rgs_gaussian* g, size_t n, const rgs_camera* cam, const float bg[3],
                     float* rgb_out);

/* rasterize (cpu/rasterizer.rs:72-192): rgba_out is h*w*4 (the Framebuffer
 * data; pixels of tiles no splat touches stay 0). */
int rgs_rasterize_rgba(rgs_ctx* ctx, const rgs_gaussian* g, size_t n, const rgs_camera* cam, const float bg[3],
                       float* rgba_out);

/* photometric_loss_grad (cpu/backward.rs:646-693): rgb is w*h*3, target
 * w*h*4, dl_drgb_out w*h*3. */
int rgs_loss_grad(rgs_ctx* ctx, const float* rgb, const float* target_rgba, uint32_t w, uint32_t h,
                  float lambda_dssim, float* loss_out, float* dl_drgb_out);

/* photometric_loss_grad on device buffers, allocating nothing. No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). rgb_dev (w*h*3), target_rgba_dev (w*h*4), dl_drgb_dev
 * (w*h*3) and loss_dev (1 float) are device memory on the context's device;
 * the work is enqueued on the context's stream (rgs_set_stream) and complete
 * in stream order, without a host synchronisation. dL/drgb and the loss
 * are those of rgs_loss_grad (the same kernels). The kernels' partial sums
 * live in a static array of 1.5 MiB per GPU: images of at most 65536 loss
 * units of 54 x 155 pixels (any image up to 16384 x 16384); larger ones
 * return RGS_E_CAPACITY. Every call on one GPU shares that
 * array, so calls that overlap in time on one GPU (other streams, other
 * contexts, other threads) race; calls in one stream order do not. */
int rgs_loss_grad_dev_noalloc(rgs_ctx* ctx, const float* rgb_dev, const float* target_rgba_dev, uint32_t w,
                              uint32_t h, float lambda_dssim, float* loss_dev, float* dl_drgb_dev);

/* adam_step (pipeline/optimizer.rs:116-163): updates g and st in place. */
int rgs_adam_step(rgs_ctx* ctx, rgs_gaussian* g, const rgs_grad* grads, size_t n, rgs_adam* st,
                  const rgs_train_cfg* cfg);

/* densify_and_prune (pipeline/optimizer.rs:206-272): g has room for
 * `capacity` Gaussians, *n is updated. Fails with RGS_E_CAPACITY if the result
 * does not fit (g and st are then unchanged). */
int rgs_densify_and_prune(rgs_ctx* ctx, rgs_gaussian* g, size_t* n, size_t capacity, rgs_adam* st,
                          const rgs_train_cfg* cfg, float scene_extent);

/* reset_opacities (pipeline/optimizer.rs:314-319). */
int rgs_reset_opacities(rgs_ctx* ctx, rgs_gaussian* g, size_t n);

/* ---- Device-pointer forms for framework bindings (the torch.autograd
 * Function over render_color / the raster backward). g_dev is n x 59 f32 in
 * Gaussian3D order on the device. rgs_forward_dev keeps the binning state in
 * the context for the next rgs_backward_dev. */
int rgs_forward_dev(rgs_ctx* ctx, const float* g_dev, size_t n, const rgs_camera* cam, const float bg[3],
                    float* rgb_dev);
/* dL/drgb (h*w*3, device) -> grads_dev (n x 60 GaussianGrad, device). */
int rgs_backward_dev(rgs_ctx* ctx, const float* dl_drgb_dev, float* grads_dev);

/* ---- Device-resident training (the step body of pipeline/mod.rs:232-247,
 * batched over views). Parameters, Adam moments and densify statistics stay
 * in HBM between calls. */
int rgs_scene_upload(rgs_ctx* ctx, const rgs_gaussian* g, size_t n, size_t capacity);
int rgs_scene_size(rgs_ctx* ctx, size_t* n);
int rgs_scene_download(rgs_ctx* ctx, rgs_gaussian* g, size_t capacity, size_t* n);
/* Downloads the Adam state (m1/m2 rows x 59, accum, denom, t) into st. */
int rgs_adam_download(rgs_ctx* ctx, rgs_adam* st);
/* Registers the training views and their targets (w*h*4 f32 each). */
int rgs_set_views(rgs_ctx* ctx, const rgs_camera* cams, const float* const* targets_rgba, int n_views);
/* As rgs_set_views from decoded RGBA8 images of the cameras' sizes (w*h*4
 * bytes each), converted to linear RGBA on the device as load_target does
 * after decoding and resizing (io/images.rs:63-69); 4x less to upload. */
int rgs_set_views_rgba8(rgs_ctx* ctx, const rgs_camera* cams, const uint8_t* const* images_rgba8, int n_views);
/* As rgs_set_views_rgba8 from decoded RGBA8 images of any size (sizes: w, h
 * per view): an image whose size differs from its camera's is resized on the
 * device as load_target does (io/images.rs:57-61, image 0.25's
 * resize_exact(.., FilterType::Triangle)) before the conversion. */
int rgs_set_views_rgba8_sized(rgs_ctx* ctx, const rgs_camera* cams, const uint8_t* const* images_rgba8,
                              const uint32_t* sizes, int n_views);
/* image 0.25's resize_exact(.., FilterType::Triangle) of an RGBA8 image
 * (imageops/sample.rs: a vertical pass to f32, then a horizontal pass that
 * rounds to bytes), on the device. */
int rgs_resize_rgba8(rgs_ctx* ctx, const uint8_t* src, uint32_t src_w, uint32_t src_h, uint8_t* dst, uint32_t dst_w,
                     uint32_t dst_h);
/* load_target's conversion (io/images.rs:63-69, srgb_to_linear :76-82) of
 * n_pixels RGBA8 pixels to linear RGBA f32, bit-identical (the sRGB table is
 * computed with the libm powf the reference calls). */
int rgs_rgba8_to_linear(rgs_ctx* ctx, const uint8_t* rgba8, size_t n_pixels, float* out_rgba);
/* Framebuffer::to_u8_srgb (cpu/rasterizer.rs:38-49) of n_pixels RGBA f32
 * pixels. linear_to_srgb (:271-279) takes powf from an f64 pow rounded to
 * f32, so a byte can differ from the reference's only where the scaled value
 * lies within an ulp of an integer. */
int rgs_to_u8_srgb(rgs_ctx* ctx, const float* rgba, size_t n_pixels, uint8_t* out_rgba8);
/* One training iteration over the views listed in view_ids: forward and
 * backward of every listed view in single launches, gradients summed over the
 * views, one Adam step. loss_out (may be NULL) receives one loss per listed
 * view. use_graph != 0 replays the step as a captured hipGraph. */
int rgs_train_step(rgs_ctx* ctx, const int* view_ids, int n_ids, const rgs_train_cfg* cfg, const float bg[3],
                   float lambda_dssim, int use_graph, float* loss_out);
/* densify_and_prune / reset_opacities on the device-resident scene. */
int rgs_scene_densify(rgs_ctx* ctx, const rgs_train_cfg* cfg, float scene_extent);
/* The split of densify (rgs_scene_densify and rgs_densify_and_prune). 0, the
 * default, is split_gaussian (optimizer.rs:274-301): the two children's
 * log_scale[0] becomes ln(max(s0 / sqrt 2, 1e-6)), the other axes keep
 * theirs. A divisor d > 1 sets every child log-scale to ln(max(s_j / d,
 * 1e-6)) -- the "scale / 1.6" split of 3DGS practice -- with the
 * reference's offsets and opacity. No counterpart in the reference. */
int rgs_set_split_scale_divisor(rgs_ctx* ctx, float divisor);
int rgs_scene_reset_opacities(rgs_ctx* ctx);
/* Last render's (tile, splat) pair count, summed over its views, and the
 * (pixel, splat) pairs its forward and backward passes evaluate (each output
 * pointer may be NULL). */
int rgs_step_stats(rgs_ctx* ctx, uint64_t* n_pairs, uint64_t* fwd_evals, uint64_t* bwd_evals);

/* ---- Depth-cut binning of rgs_train_step (no reference counterpart; images,
 * losses and gradients do not depend on it). A tile's list only matters up to
 * the splat at which its last pixel saturates (T < 1e-4), so the step bins
 * and sorts the pairs of the splats in front of a cut depth first, and only
 * the tiles those lists leave unsaturated get the pairs behind the cut.
 * mode 0 (default): the cut is the previous step's saturation depth + 5%
 * (none on the first step), paused for 256 steps (whole lists) after a step
 * that needed back lists while fewer than an eighth as many visible splats
 * lay behind the cut as in front of it, and resumed with the back-list stage
 * on its first two steps; 1: no cut (whole lists); 2: cut at the fixed
 * camera-space depth `depth` (> 0). While the previous step needed no back
 * list, a step runs without the back-list stage (a captured graph of its
 * own); if it then leaves a bin unsaturated, the device skips its update and
 * the step is re-run with the stage. rgs_cut_stats reports the last step's
 * front and back pair counts, how many workgroups needed back lists, the cut
 * depth the next step will use (+inf: none) and how many steps were re-run
 * so far; pointers may be NULL. */
int rgs_set_depth_cut(rgs_ctx* ctx, int mode, float depth);
int rgs_cut_stats(rgs_ctx* ctx, uint64_t* front_pairs, uint64_t* back_pairs, uint64_t* flagged_tiles,
                  float* next_cut, uint64_t* reruns);
/* The images the last rgs_train_step rendered (its views in order, h x w x 3
 * each; n_floats >= 3 x total pixels). */
int rgs_step_image(rgs_ctx* ctx, float* rgb, size_t n_floats);

/* ---- Raster work decomposition (no reference counterpart; images do not
 * depend on it, gradients only through the order of float summation).
 * super_tile: side, in 16x16 tiles, of the square patches the raster
 * schedule walks (1..1024; large = row-major), which sets what the blocks
 * resident on one XCD share in L2; 0: row-major, except the training
 * step's backward, which takes its bins longest first (in descending order
 * of how far down their lists the step's forward composited), so that the
 * launch ends on short bins. sparse_lanes: in the backward, a splat
 * whose gradient reaches at most this many lanes of a wave is accumulated
 * with LDS atomics instead of a cross-lane reduction (0..64).
 * rgs_autotune_raster renders the registered views ids on the device-resident
 * scene and times the forward + backward raster under candidate settings
 * (coordinate search over super_tile, sparse_lanes and, with binning, the
 * staging batch; best of two runs each), keeps the fastest and reports it
 * (the batch through rgs_get_raster_batch); the scene and optimizer state
 * are not changed. */
int rgs_set_raster_params(rgs_ctx* ctx, int super_tile, int sparse_lanes);
/* Splats the binned training raster stages in LDS per batch: 32, 64 (the
 * default) or 128 (no reference counterpart; images do not depend on it,
 * gradients only through the order of float summation). */
int rgs_set_raster_batch(rgs_ctx* ctx, int batch);
int rgs_get_raster_batch(rgs_ctx* ctx, int* batch);
/* Persistent tile queue of the training step's raster (no reference
 * counterpart; images do not depend on it, gradients only through the order
 * of float summation). Bit 0: the forward, bit 1: the backward run as
 * persistent kernels, a grid of as many workgroups as the device holds at
 * once, each taking bins from per-XCD queues (its own XCD's share of the
 * schedule first, then the others') until none is left; 0 (the default):
 * one workgroup per bin. rgs_autotune_raster searches it with the binned
 * training raster. */
int rgs_set_raster_queue(rgs_ctx* ctx, int mode);
int rgs_get_raster_queue(rgs_ctx* ctx, int* mode);
/* Binning of rgs_train_step: pairs per bin of 2^shift x 2^shift tiles
 * (0: the reference's 16x16 tiles; 1, the default: 32x32). The backward
 * rasterises a bin with one workgroup of four waves that share the staged
 * splats and merge each splat's gradient before the accumulator atomics;
 * the forward does the same with fwd_tile_blocks = 0, or runs one
 * single-wave workgroup per tile of the bin (1, the default). API renders
 * always use the reference's tiles. */
int rgs_set_train_binning(rgs_ctx* ctx, int shift, int fwd_tile_blocks);
int rgs_get_raster_params(rgs_ctx* ctx, int* super_tile, int* sparse_lanes);
int rgs_autotune_raster(rgs_ctx* ctx, const int* view_ids, int n_ids, const float bg[3], float lambda_dssim,
                        int* super_tile, int* sparse_lanes, float* raster_ms);

/* ---- Preview renderer (gpu/compute.rs:138-297, gpu/splat.wgsl:120-190).
 * GpuRenderer's API: project gives n SplatInfo2D records of 10 floats (sx,
 * sy, conic a, b, c, colour r, g, b, alpha, depth; all zero where the shader
 * culls), render gives linear RGB h*w*3: DC colour only, the shader's
 * covariance (no perspective terms) and composite's 3-sigma rects, nearest
 * first (equal depths in index order). rgs_scene_preview_render renders the
 * device scene (a viewer during training, no host copy of the scene).
 * WGSL math runs as IEEE f32 operations in source order. */
int rgs_preview_project(rgs_ctx* ctx, const rgs_gaussian* g, size_t n, const rgs_camera* cam, float* splats_out);
int rgs_preview_render(rgs_ctx* ctx, const rgs_gaussian* g, size_t n, const rgs_camera* cam, const float bg[3],
                       float* rgb_out);
int rgs_scene_preview_render(rgs_ctx* ctx, const rgs_camera* cam, const float bg[3], float* rgb_out);
/* As rgs_scene_preview_render into device memory (h*w*3 floats on the
 * context's device, written in stream order on the context's stream): a
 * viewer's frame without the copy to the host. */
int rgs_scene_preview_render_device(rgs_ctx* ctx, const rgs_camera* cam, const float bg[3], float* rgb_dev);

/* ---- SfM initialisation (pipeline/sfm.rs:365-455). gaussians_from_sfm for
 * n points (xyz: n x 3 f64, rgb: n x 3 bytes; SfMPoint's other fields are
 * not used): positions, sRGB-to-linear DC colours, identity rotations,
 * opacity logit(0.1) and isotropic log-scales from the mean distance to the
 * 3 nearest of a 50-point window in x order (estimate_initial_scales). The
 * log is the f64 log rounded to f32 (within 1 ulp of the reference's logf);
 * every other value is the reference's. RGS_E_NAN where the reference panics
 * (a NaN x, or a NaN distance in a window). rgs_scene_from_sfm builds the
 * device scene directly (as rgs_scene_upload). */
int rgs_gaussians_from_sfm(rgs_ctx* ctx, const double* xyz, const uint8_t* rgb, size_t n, rgs_gaussian* out);
int rgs_scene_from_sfm(rgs_ctx* ctx, const double* xyz, const uint8_t* rgb, size_t n, size_t capacity);

/* ---- Model files (io/ply.rs). The records move between the device
 * parameter rows and the file as one block; headers are text on the host.
 * rgs_save_ply replaces save_ply (io/ply.rs:35-102): binary little-endian
 * 3DGS layout, byte-identical to the reference's file. rgs_scene_save_ply
 * writes the device scene (after rgs_scene_upload / training) without a host
 * copy of the records in Gaussian3D layout.
 * rgs_load_ply replaces load_ply (io/ply.rs:116-247): binary (either
 * endianness) or ASCII, any property order, float/double/uchar/ushort/int
 * properties decoded as the reference does, Gaussian3D::new defaults for
 * absent properties. With out == NULL or capacity < count it returns
 * RGS_E_CAPACITY and *n_out = count. rgs_scene_load_ply loads into the device
 * scene (as rgs_scene_upload: moments and statistics cleared).
 * rgs_save_splat replaces save_splat (io/ply.rs:258-303): 32-byte records. */
int rgs_save_ply(rgs_ctx* ctx, const rgs_gaussian* g, size_t n, const char* path);
int rgs_scene_save_ply(rgs_ctx* ctx, const char* path);
int rgs_load_ply(rgs_ctx* ctx, const char* path, rgs_gaussian* out, size_t capacity, size_t* n_out);
int rgs_scene_load_ply(rgs_ctx* ctx, const char* path, size_t capacity);
int rgs_save_splat(rgs_ctx* ctx, const rgs_gaussian* g, size_t n, const char* path);
int rgs_scene_save_splat(rgs_ctx* ctx, const char* path);

/* ---- Stage timing (no reference counterpart). With profiling on,
 * rgs_train_step brackets each stage with HIP events on the context's stream;
 * in graph mode the events are nodes of the captured graph, so replays are
 * timed as they run. Stages, in order: 0 project, 1 depth sort,
 * 2 pair-count scan, 3 pair generation, 4 tile sort (its last pass also
 * emits the tile ranges), 5 forward raster, 6 L1+D-SSIM, 7 backward raster,
 * 8 geometry + Adam, 9 back lists of the depth cut (count, scan, generation,
 * sort, resumed forward). Under the depth cut, stages 3-5 cover the front
 * pairs (and 3 the cut itself).
 * rgs_set_profiling_stages brackets only the stages whose bit is set in
 * stage_mask (bit s = stage s; 0 turns profiling off), so a timed run can
 * carry just the events it reads; rgs_set_profiling(ctx, enable) sets all or
 * none.
 * rgs_stage_times returns the mean milliseconds per stage since profiling
 * was enabled (0 for stages outside the mask), the number of profiled steps
 * and how many of them were graph replays. */
#define RGS_N_STAGES 10
int rgs_set_profiling(rgs_ctx* ctx, int enable);
int rgs_set_profiling_stages(rgs_ctx* ctx, uint32_t stage_mask);
int rgs_stage_times(rgs_ctx* ctx, float* mean_ms, int n_stages, uint64_t* steps, uint64_t* graph_steps);

/* ---- Diagnostics (no reference counterpart; used by the parity tests).
 * rgs_eval_expf evaluates the device expf every kernel uses on n host
 * values. rgs_debug_tiles returns the last render's tile lists as CSR of
 * Gaussian indices in list order (offsets has total_tiles + 1 entries; gis
 * room for k_max entries; *k_out receives the pair count). */
int rgs_eval_expf(rgs_ctx* ctx, const float* x, float* out, size_t n);
int rgs_debug_tiles(rgs_ctx* ctx, uint64_t* offsets, size_t n_offsets, uint32_t* gis, size_t k_max,
                    uint64_t* k_out);
/* Gradient capture of rgs_train_step: with enable != 0, each step also
 * writes the gradients its Adam update consumed (summed over the step's
 * views, GaussianGrad order; screen_grad_norm is the sum of the views'
 * norms), and rgs_step_grads returns the last step's for the n = scene size
 * Gaussians. Replaces nothing in the reference: it exposes the values the
 * fused geometry/Adam kernel keeps in registers, so that the step's
 * gradients and its Adam update can each be compared with the CPU path. */
int rgs_set_grad_capture(rgs_ctx* ctx, int enable);
/* The work the last rgs_train_step's raster kernels did, counted by running
 * their counting instantiations on the same lists (images, losses and the
 * scene are unchanged): counts[0] forward lane evaluations of a (pixel,
 * splat) pair, [1] forward wave visits of a (splat, 4-row pixel group) that
 * its alpha-reach test kept, [2] forward composites (alpha >= 1/255), [3]-[5]
 * the same for the backward (contributing pairs in [5]). Dividing a
 * profiler's VALU instruction count by [1] or [4] gives instructions per
 * visit. Not for a step whose depth cut resumed tiles on back lists. */
int rgs_raster_work_counts(rgs_ctx* ctx, uint64_t* counts);
int rgs_step_grads(rgs_ctx* ctx, rgs_grad* grads_out, size_t n);
/* Measurement aid (no reference counterpart): the HBM bandwidth a plain
 * streaming copy reaches on the context's device, the practical ceiling of
 * the memory-bound kernels. Copies `bytes` (rounded down to 16 B, at most
 * 1 TiB) from one fresh device buffer to another, one 16-byte load and store
 * per thread, `reps` timed
 * launches after two untimed ones, and returns read + written bytes per
 * second of the mean launch in *gbs (GB/s). Allocates and frees its two
 * buffers. */
int rgs_hbm_copy_probe(rgs_ctx* ctx, size_t bytes, int reps, float* gbs);

#ifdef __cplusplus
}
#endif
#endif /* RGSPLAT_HIP_H */
