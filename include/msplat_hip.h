/*
 * msplat HIP backend — C ABI.
 *
 * The drop-in boundary for msplat's backend interface core/metal/bindings.h (v1.1.4):
 * every entry point below names the bindings.h function it replaces. A C++ backend file
 * implementing bindings.h over MTensor (core/hip/bindings_hip.cpp) or any FFI
 * (ctypes, cgo, JNI) calls these with plain device pointers and sizes.
 *
 * Conventions
 *  - Pointers documented "device" are HIP device allocations (msplat_hip_malloc or any
 *    other hipMalloc'ed memory); "host" pointers are read during the call.
 *  - Gaussian parameters use the reference layouts (model.cpp:52-64): means (N,3),
 *    log-scales (N,3), quaternions (N,4) stored (w,x,y,z), features_dc (N,3),
 *    features_rest (N, bases-1, 3), opacity logits (N) — all float32, row-major.
 *  - Every call is enqueued on the context's stream and returns without waiting unless
 *    it says "syncs". Outputs alias per-context caches where the reference's do
 *    (bindings.h:93-96) and stay valid until the next call on the same context.
 *  - Every function returns MSPLAT_OK (0) or an error code; msplat_hip_last_error()
 *    gives the message. No C++ exception crosses this boundary.
 */
#ifndef MSPLAT_HIP_H
#define MSPLAT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MSPLAT_OK = 0,
    MSPLAT_ERR_ARG = 1,      /* invalid argument (std::invalid_argument in the reference) */
    MSPLAT_ERR_HIP = 2,      /* a HIP runtime call failed */
    MSPLAT_ERR_RANGE = 3,    /* a size limit was exceeded (e.g. > 2^31 intersections, cpu:546-547) */
    MSPLAT_ERR_CAPACITY = 4  /* densify with 3N > capacity (cpu:1347) */
};

typedef struct msplat_hip_ctx msplat_hip_ctx;

/* One view: the values Model::prepareCam computes (model.cpp:498-546). */
typedef struct {
    float viewmat[16]; /* world -> camera (OpenCV axes), row-major */
    float projmat[16]; /* projection * viewmat, row-major */
    float cam_pos[3];
    float fx, fy, cx, cy;
    uint32_t width, height;
    int32_t tiles_x, tiles_y; /* ceil(W/16), ceil(H/16) */
    float clip_thresh;        /* 0.01 in Model */
    float glob_scale;         /* 1.0 in Model */
    uint32_t degree;          /* SH degree the rest coefficients are stored for */
    uint32_t degrees_to_use;  /* SH degree evaluated this step */
} msplat_hip_camera;

/* Device pointers to the six parameter groups. */
typedef struct {
    int32_t num_points;
    float *means, *scales, *quats, *features_dc, *features_rest, *opacities;
} msplat_hip_gaussians;

/* The six Adam groups (means, scales, quats, dc, rest, opacity), as msplat_train_step
 * takes them (bindings.h:100-118). params[g] must equal the matching gaussians pointer. */
typedef struct {
    float *exp_avg[6], *exp_avg_sq[6];
    float step_size[6]; /* lr / (1 - beta1^t) */
    float bc2_sqrt[6];  /* sqrt(1 - beta2^t) */
    float beta1, beta2, eps;
} msplat_hip_adam;

/* Densification statistics, accumulated by the train step (cpu:1311-1324). */
typedef struct {
    float *vis_counts, *xys_grad_norm, *max_2d_size; /* device, (N) */
} msplat_hip_grad_stats;

/* ── context, memory, streams (bindings.h:7-29) ─────────────────────────────── */

const char *msplat_hip_last_error(void);

/* Creates a context on `device` using `stream` (a hipStream_t; NULL creates a private
 * stream). Replaces the global Metal context and FusedTensorCache (msplat_metal.mm:336-585). */
int msplat_hip_ctx_create(int device, void *stream, msplat_hip_ctx **out);
/* cleanup_msplat_metal (bindings.h:8): releases every cached buffer. */
int msplat_hip_ctx_destroy(msplat_hip_ctx *ctx);
int msplat_hip_ctx_set_stream(msplat_hip_ctx *ctx, void *stream);
void *msplat_hip_ctx_stream(msplat_hip_ctx *ctx);
/* msplat_gpu_sync (bindings.h:26): waits for everything enqueued on the context. */
int msplat_hip_sync(msplat_hip_ctx *ctx);
/* msplat_commit (bindings.h:23): nothing to commit on HIP; kept for the interface. */
int msplat_hip_commit(msplat_hip_ctx *ctx);

/* gpu_empty / gpu_zeros (bindings.h:18-19): device allocation. */
int msplat_hip_malloc(size_t bytes, int zero, void **out);
int msplat_hip_free(void *ptr);
/* Free and total device memory of `device` (no reference counterpart: the serving
 * loop's default HBM budget). */
int msplat_hip_device_memory(int device, size_t *free_bytes, size_t *total_bytes);
/* Copies on the context's stream; kind 0 = host->device, 1 = device->host (syncs),
 * 2 = device->device (msplat_copy_buffer, bindings.h:143). */
int msplat_hip_memcpy(msplat_hip_ctx *ctx, void *dst, const void *src, size_t bytes, int kind);
int msplat_hip_memset(msplat_hip_ctx *ctx, void *dst, int value, size_t bytes);

/* ── render and train (bindings.h:81-118) ───────────────────────────────────── */

/* msplat_render (bindings.h:81-91): projection + SH, (tile|depth) sort, forward
 * compositing. background: host float[3]. *out_img receives the context's cached
 * (H, W, 3) image (device), valid until the next call. */
int msplat_hip_render(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_camera *cam,
                      const float background[3], float **out_img);

/* msplat_train_step (bindings.h:100-118) without the prior-guided extras: forward,
 * L1 + SSIM loss (ssim_weight, loss_inv_n = 1/(3HW)), backward, Adam on all six groups
 * (SH groups fused into the projection backward), densification statistics.
 * gt: device (H, W, 3). radii_out (device, N int32) and loss_out (host; syncs) may be
 * NULL. The loss is also kept on the device (msplat_hip_last_loss). */
int msplat_hip_train_step(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_camera *cam,
                          const float background[3], const float *gt, float ssim_weight, float loss_inv_n,
                          const msplat_hip_adam *adam, const msplat_hip_grad_stats *stats, float inv_max_dim,
                          int32_t *radii_out, float *loss_out);

/* The loss of the last train step (host; syncs). */
int msplat_hip_last_loss(msplat_hip_ctx *ctx, float *loss);

/* ── prior-guided training (PriorStep, bindings.h:41-77) ─────────────────────
 * The fields of PriorStep with device pointers for its tensors. aux rasterizes the
 * expected-depth sum and alpha over a per-pixel background (the learned sky, else the
 * constant) and is required by the sky and by every prior loss; sky: equirect texture
 * (sky_h, sky_w, 3) seen by view direction in the frame sky_frame (up, azimuth 0, azimuth
 * 90 deg), with its gradient and Adam moments (train only); prior_depth (prior_h,
 * prior_w) float (0 = invalid) and prior_aux (prior_h, prior_w, 4) uint8 (confidence,
 * sky, keep, -) at the prior's own resolution; loss_terms [4] float: depth, sky, fill, 0
 * (rewritten every step); exposure: per-camera affine colour rows expo_params (num_cams,
 * 12) with their moments and an extra gradient expo_grad [12] (cleared after use);
 * log_max_scale_ratio > 0 caps each gaussian's largest scale at exp(ratio) x its middle
 * one after the Adam step. A prior-guided frame is rasterized one workgroup per tile. */
typedef struct {
    int32_t aux, sky;
    float *sky_tex, *sky_grad, *sky_exp_avg, *sky_exp_avg_sq;
    int32_t sky_w, sky_h;
    float sky_frame[9];
    float sky_step_size, sky_bc2_sqrt;
    const float *prior_depth;
    const uint8_t *prior_aux;
    int32_t prior_w, prior_h, has_depth, has_sky_mask;
    float depth_weight, sky_weight, fill_weight, huber_delta, min_alpha;
    int32_t mask_photometric;
    float *loss_terms;
    int32_t exposure, cam_index, num_cams;
    float *expo_params, *expo_exp_avg, *expo_exp_avg_sq, *expo_grad;
    float expo_step_size, expo_bc2_sqrt, expo_reg;
    float log_max_scale_ratio;
} msplat_hip_prior;

/* msplat_train_step with a PriorStep (bindings.h:100-118, cpu:1204-1327): the aux
 * forward, the photometric loss on the exposure-adjusted render, the prior losses (which
 * mask the photometric gradient where the prior does not keep a pixel), the exposure and
 * sky backward with their Adam steps, the aux raster backward (depth gradients move the
 * means along the view axis), Adam on the six groups, the needle cap, the statistics.
 * Other arguments as msplat_hip_train_step; loss_out is the photometric loss. */
int msplat_hip_train_step_prior(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_camera *cam,
                                const float background[3], const float *gt, float ssim_weight, float loss_inv_n,
                                const msplat_hip_adam *adam, const msplat_hip_grad_stats *stats, float inv_max_dim,
                                int32_t *radii_out, float *loss_out, const msplat_hip_prior *prior);
/* msplat_render with a PriorStep (bindings.h:81-91): with aux, the learned sky behind
 * the gaussians and the aux outputs. */
int msplat_hip_render_prior(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_camera *cam,
                            const float background[3], const msplat_hip_prior *prior, float **out_img);
/* msplat_render_aux_outputs (bindings.h:93-96): after an aux frame, the expected-depth sum
 * sum_i w_i z_i and the final transmittance (device (H, W), the context's caches). */
int msplat_hip_render_aux_outputs(msplat_hip_ctx *ctx, float **depth, float **final_T);
/* msplat_finalize_depth (bindings.h:177-180): alpha = 1 - final_T and the expected depth
 * depth_num / alpha * inv_scale (0 where alpha <= 1e-4), n pixels, device arrays. */
int msplat_hip_finalize_depth(msplat_hip_ctx *ctx, const float *depth_num, const float *final_T, uint32_t n,
                              float inv_scale, float *depth_out, float *alpha_out);

/* ── differentiable rendering (new: torch.autograd.Function backend) ─────────
 * render_backward takes dL/d(image) (device (H, W, 3)) of the last msplat_hip_render
 * on this context and writes the parameter gradients (device, reference layouts). The
 * view direction of the SH colour is not differentiated (as in the reference,
 * cpu:1111-1137). Any output pointer may be NULL. */
typedef struct {
    float *v_means, *v_scales, *v_quats, *v_features_dc, *v_features_rest, *v_opacities;
} msplat_hip_grads;
int msplat_hip_render_backward(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_camera *cam,
                               const float background[3], const float *v_img, const msplat_hip_grads *out);

/* ── the per-frame cache (the reference's Frame / FusedTensorCache) ──────────
 * Device pointers to the intermediate buffers of the last render/train call, for tests
 * and tools. Per gaussian: xys (N,2) as float4 records (x, y, conic.x, conic.y) and
 * (conic.z, sigmoid(opacity), r, g) plus blue (N), depths, radii, aabb (N,2), tile
 * counts, prefix offsets. Per intersection (sorted): keys (tile<<32 | depth bits,
 * filled by frame_view when MSPLAT_FRAME_KEYS64 is set, else NULL), gaussian ids (NULL
 * without MSPLAT_FRAME_KEYS64 when MSPLAT_TUNE_PACK_KEYS packed them into the sort keys).
 * tile_start (T+1): tile t owns sorted intersections [tile_start[t], tile_start[t+1]).
 * Per pixel: image, final_T, final_idx (index into
 * the sorted intersections), dL/d(image). Per gaussian reduced gradients (N,2)/(N,3)/
 * (N,3)/(N) are filled when MSPLAT_FRAME_KEEP_GRADS is set. loss_sum: the last loss's
 * double sum of the per-pixel losses (one double, device memory). slots_written: a
 * running count (one uint64, device memory) of the gradient slots the one-view raster
 * backward has written since the context was created; the other slots of its frames
 * were all-zero and only flagged. */
typedef struct {
    int32_t num_points, num_tiles;
    int64_t num_isects;
    float *splat_a, *splat_b, *splat_blue, *depths, *aabb;
    int32_t *radii;
    uint32_t *tile_count, *offsets;
    uint64_t *isect_keys;
    int32_t *isect_gid;
    int32_t *tile_start;
    float *image, *final_T, *v_image;
    int32_t *final_idx;
    float *v_xy, *v_conic, *v_rgb, *v_opacity;
    double *loss_sum;
    uint64_t *slots_written;
    int64_t num_visible; /* gaussians with a non-empty tile rectangle (radius > 0) */
    int32_t depth_chunks; /* depth chunks per tile of the last one-view forward (1: none) */
    /* device pointers to running counts (uint64) of the one-view frames' intersections and
     * visible gaussians since the context's first frame (a replay run again after an
     * intersection overflow counts twice); read before and after a window of steps */
    uint64_t *isects_running;
    uint64_t *visible_running;
} msplat_hip_frame;
int msplat_hip_frame_view(msplat_hip_ctx *ctx, msplat_hip_frame *out);
/* KEEP_GRADS: fill the reduced gradients; KEYS64: materialise isect_keys on frame_view
 * (and isect_gid, when the tile sort packed the ids into its keys). */
enum { MSPLAT_FRAME_KEEP_GRADS = 1, MSPLAT_FRAME_KEYS64 = 2 };
int msplat_hip_set_flags(msplat_hip_ctx *ctx, int flags);

/* Kernel variants picked per context (by the autotuner, msplat/autotune.py); no
 * reference counterpart (the Metal backend hard-codes its launch shapes,
 * msplat_metal.mm:431-436). Every variant computes the same result.
 *   MSPLAT_TUNE_BWD_PPT    backward rasterizer pixels per thread: 1, 2, 4 (default 4)
 *   MSPLAT_TUNE_FWD_BATCH  forward LDS batch depth (gaussian records staged per batch):
 *                          64, 128, 256 (default 256)
 *   MSPLAT_TUNE_BWD_BATCH  backward LDS batch depth: 32, 64, 128 (default 64)
 *   MSPLAT_TUNE_CULL       1 (default): the rasterizers skip a tile's gaussians whose
 *                          alpha >= 1/255 ellipse (widened for float rounding) misses
 *                          the tile; 0: every listed gaussian is visited. Skipped pairs
 *                          would composite nothing, so both give identical results.
 *   MSPLAT_TUNE_PACK_KEYS  1 (default): when the tile-id and gaussian-id bits fit in 32,
 *                          the one-view tile sort moves (tile << id bits | id) keys alone
 *                          instead of tile keys plus id values (same order, half the
 *                          bytes); 0: always keys plus values.
 *   MSPLAT_TUNE_CHUNKS     depth chunks per tile of the one-view rasterizers (the depth-
 *                          chunked form of msplat_metal.mm:782-907 for views with few
 *                          tiles): 0 (default) picks per frame (fewer than 400 tiles:
 *                          chunks of at least 16 intersections on average, up to about
 *                          16384 workgroups, 2 to 64 chunks; else none), 1 never chunks,
 *                          2..64 fixes the count. Chunked frames equal the one-workgroup-
 *                          per-tile result up to float rounding of the transmittance
 *                          products and colour sums, not bit for bit. The batched
 *                          rasterizers do not chunk.
 * The binning tile stays 16x16: it decides which gaussians reach a pixel. */
enum {
    MSPLAT_TUNE_BWD_PPT = 0,
    MSPLAT_TUNE_FWD_BATCH = 1,
    MSPLAT_TUNE_BWD_BATCH = 2,
    MSPLAT_TUNE_CULL = 3,
    MSPLAT_TUNE_PACK_KEYS = 4,
    MSPLAT_TUNE_CHUNKS = 5,
    MSPLAT_NUM_TUNABLES = 6
};
int msplat_hip_set_tuning(msplat_hip_ctx *ctx, int key, int value);
int msplat_hip_get_tuning(msplat_hip_ctx *ctx, int key, int *value);

/* Stage-wise entry points over the frame cache (test hooks, like the reference's
 * _core._gpu_* hooks, python/bindings.cpp:655-738). Each runs one stage of
 * msplat_render / msplat_train_step on the context's stream. */
int msplat_hip_stage_project(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_camera *cam);
/* Binning from the frame's projection outputs (which a caller may overwrite first). */
int msplat_hip_stage_bin(msplat_hip_ctx *ctx, const msplat_hip_camera *cam);
int msplat_hip_stage_raster_forward(msplat_hip_ctx *ctx, const msplat_hip_camera *cam, const float background[3]);
/* L1 + SSIM loss of the frame image against gt; fills v_image, returns the loss (syncs). */
int msplat_hip_stage_loss(msplat_hip_ctx *ctx, const msplat_hip_camera *cam, const float *gt, float ssim_weight,
                          float loss_inv_n, float *loss_out);

/* ── stage timing (msplat_enable_gpu_timing / msplat_drain_stage_times,
 * bindings.h:34-41) ─────────────────────────────────────────────────────────
 * HIP events on the context's stream around each stage of render/train/backward:
 * 0 proj_sh_fwd, 1 prefix_sort_pack, 2 rast_fwd, 3 loss_fwd_bwd, 4 rast_bwd,
 * 5 proj_sh_bwd_adam. drain syncs, writes the accumulated milliseconds and launch
 * counts per stage (max_stages entries each) and clears them. */
enum { MSPLAT_NUM_STAGES = 6 };
int msplat_hip_enable_stage_timing(msplat_hip_ctx *ctx, int enable);
int msplat_hip_drain_stage_times(msplat_hip_ctx *ctx, double *ms, int64_t *calls, int max_stages, int *n_stages);
/* Times only the stages whose bit (1 << stage) is set in mask (default: every
 * stage). Each timed stage records two events, and an event between two kernels
 * costs the stream some microseconds, so a caller measuring whole-step throughput
 * times the one stage it needs. No counterpart in the reference. */
int msplat_hip_set_stage_mask(msplat_hip_ctx *ctx, uint32_t mask);
const char *msplat_hip_stage_name(int stage);

/* ── densification and maintenance (bindings.h:120-147) ─────────────────────── */

/* msplat_densify (bindings.h:120-135): classify, split (2 children, PCG normals),
 * duplicate, cull, stable in-place compaction of the 18 buffers (params, exp_avg,
 * exp_avg_sq), each with `capacity` rows. Syncs (returns the new count in *new_count). */
int msplat_hip_densify(msplat_hip_ctx *ctx, int32_t n, int32_t capacity, float grad_thresh, float size_thresh,
                       float screen_thresh, int check_screen, float cull_alpha_thresh, float cull_scale_thresh,
                       float cull_screen_size, int check_huge, const float *xys_grad_norm, const float *vis_counts,
                       const float *max_2d_size, float half_max_dim, float *const bufs[18], int32_t fr_stride,
                       uint32_t seed, int32_t *new_count);

/* msplat_opacity_reset (bindings.h:137-140). */
int msplat_hip_opacity_reset(msplat_hip_ctx *ctx, float *opacities, float *exp_avg, float *exp_avg_sq, int32_t n,
                             float reset_logit);

/* msplat_radix_sort (bindings.h:145-147): stable LSD sort of keys (u64, device) with
 * int32 payload (device) by the low key_bits bits, in place. */
int msplat_hip_radix_sort(msplat_hip_ctx *ctx, uint64_t *keys, int32_t *vals, uint32_t n, int key_bits);

/* No counterpart in the reference. For callers that cannot allocate device memory
 * (code under stream capture, tight serving loops, harnesses that forbid hipMalloc):
 * the same sort as msplat_hip_radix_sort, with identical results, that allocates
 * nothing. Its scratch is a static device array sized for MSPLAT_NOALLOC_SORT_MAX
 * elements (about 13.6 MB, held by every GPU the library runs on until the process
 * ends); larger n returns MSPLAT_ERR_RANGE before any launch. All calls on one GPU
 * share that array: calls that overlap on one GPU race. */
#define MSPLAT_NOALLOC_SORT_MAX (1u << 20)
int msplat_hip_radix_sort_noalloc(msplat_hip_ctx *ctx, uint64_t *keys, int32_t *vals, uint32_t n, int key_bits);

/* msplat_knn3_mean_dist (bindings.h:149-151): exact mean distance to the 3 nearest
 * other points. points (n,3) and mean_dist (n) are device pointers. */
int msplat_hip_knn3_mean_dist(msplat_hip_ctx *ctx, const float *points, uint32_t n, float *mean_dist);

/* msplat_init_gaussians (bindings.h:153-156). */
int msplat_hip_init_gaussians(msplat_hip_ctx *ctx, const float *means, const uint8_t *rgb, uint32_t n, uint32_t seed,
                              float opacity_logit, float *scales, float *quats, float *features_dc, float *opacities);

/* msplat_resize_area (bindings.h:160-163): area resampling of a (sh, sw) image into
 * (dh, dw, 3) float RGB (device). src is RGBA8 (4 bytes per pixel, alpha ignored, scaled
 * by 1/255) when src_is_u8, else float RGB. The ground-truth pyramid's step. */
int msplat_hip_resize_area(msplat_hip_ctx *ctx, const void *src, int src_is_u8, int sw, int sh, float *dst, int dw,
                           int dh);

/* msplat_undistort (bindings.h:165-168): Brown-Conrady undistortion of (sh, sw, 3) float
 * RGB (device) into the (dh, dw, 3) crop at (roi_x, roi_y) of the undistorted image.
 * intr = fx, fy, cx, cy; dist = k1, k2, p1, p2, k3 (host). */
int msplat_hip_undistort(msplat_hip_ctx *ctx, const float *src, int sw, int sh, const float intr[4],
                         const float dist[5], int roi_x, int roi_y, float *dst, int dw, int dh);

/* ── trained-scene files (core/src/loaders/save_gaussians.cpp, model.cpp:282-496) ──
 * Host I/O over device parameters (each call syncs the context). rest_bases is the
 * number of SH rest bases stored per gaussian ((degree+1)^2 - 1). keep_crs applies the
 * inverse of the dataset's scale and translation, as GaussianParams does. */
/* saveGaussianPly (save_gaussians.cpp:10-52): binary little-endian PLY. */
int msplat_hip_save_ply(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, int rest_bases, const char *path,
                        int step, float scale, const float translation[3], int keep_crs);
/* saveGaussianSplat (save_gaussians.cpp:54-104): 32-byte records, largest first. */
int msplat_hip_save_splat(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const char *path, float scale,
                          const float translation[3], int keep_crs);
/* loadGaussianPly's header (save_gaussians.cpp:106-131): vertex count, rest bases, step. */
int msplat_hip_ply_info(const char *path, int32_t *num_points, int32_t *rest_bases, int32_t *step);
/* loadGaussianPly (save_gaussians.cpp:106-191) into device buffers of g->num_points rows
 * (the count and rest bases must be the file's, from msplat_hip_ply_info). */
int msplat_hip_load_ply(msplat_hip_ctx *ctx, const char *path, const msplat_hip_gaussians *g, int rest_bases,
                        float scale, const float translation[3], int keep_crs);

/* Model::saveCheckpoint / loadCheckpoint (model.cpp:282-496): the .msplat v2 file
 * ("MSPL", version 2): step, gaussian count, SH degree, Adam step count and learning
 * rates, the six parameter groups and both Adam moments as shaped tensors, then the
 * learned-sky and exposure state (written absent; skipped on load, flagged in has_sky /
 * has_exposure). Device pointers; exp_avg[4] / exp_avg_sq[4] may be NULL at degree 0. */
typedef struct {
    int32_t version, step, sh_degree, adam_step_count;
    float adam_lr[6], means_lr_init, means_lr_final;
    msplat_hip_gaussians gaussians;
    float *exp_avg[6], *exp_avg_sq[6];
    int32_t has_sky, has_exposure;
} msplat_hip_checkpoint;
int msplat_hip_save_checkpoint(msplat_hip_ctx *ctx, const char *path, const msplat_hip_checkpoint *ckpt);
/* The scalar header: version, step, gaussian count (gaussians.num_points), SH degree,
 * Adam step count and learning rates; no device access. */
int msplat_hip_checkpoint_info(const char *path, msplat_hip_checkpoint *out);
/* Reads the tensors into ckpt's device buffers (num_points and sh_degree set to the
 * file's) and fills the scalar fields. */
int msplat_hip_load_checkpoint(msplat_hip_ctx *ctx, const char *path, msplat_hip_checkpoint *ckpt);

/* msplat_image_metrics (bindings.h:170-172): PSNR, SSIM (clamp-to-edge), L1. Syncs. */
int msplat_hip_image_metrics(msplat_hip_ctx *ctx, const float *rendered, const float *gt, int h, int w,
                             double out[3]);

/* msplat_pack_rgba8 (bindings.h:174-175): out is a device (n,4) uint8 buffer. */
int msplat_hip_pack_rgba8(msplat_hip_ctx *ctx, const float *img, uint32_t n, uint8_t *out);

/* ── batched views (no reference counterpart: the reference renders and trains one view
 * per call, msplat_render / msplat_train_step, bindings.h:81-118) ──────────────────
 * Several views of one scene, of mixed resolution, in one pass: one projection of every
 * gaussian into every view, one key sort over all views' tiles, a persistent rasterizer
 * that pulls (view, tile) items from a device work queue, one loss pass and one
 * backward that sums the views' gradients. Per view the results equal those of the
 * one-view entry points (images, final T and contributor indices bit for bit;
 * gradients within float rounding of the sum of one-view gradients).
 *
 * Per-pixel outputs of view v are packed at pixel offset pixel_offsets[v] (the sum of
 * W*H over the views before it): image (3 floats per pixel) and final_T. The batch
 * calls sync before they return. */
#define MSPLAT_MAX_BATCH_VIEWS 64

typedef struct {
    int32_t num_views;                 /* 1..MSPLAT_MAX_BATCH_VIEWS */
    const msplat_hip_camera *cameras;  /* host array [num_views]; same degree and glob_scale */
    const float *backgrounds;          /* host [num_views * 3] */
    const float *const *targets;       /* host array of device pointers (H_v, W_v, 3); training only */
} msplat_hip_views;

/* Renders every view; *out_images (device) receives the packed images and
 * pixel_offsets (host, num_views entries, may be NULL) the views' offsets in pixels. */
int msplat_hip_render_batch(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_views *views,
                            float **out_images, uint64_t *pixel_offsets);

/* One training step over every view: loss_v = (1 - w) L1 + w (1 - SSIM) of view v
 * (normalised by 3 W_v H_v), gradients summed over the views, one Adam step (dense
 * groups over every gaussian, SH groups over gaussians visible in some view), and the
 * densification statistics updated per view (max_2d_size with 1 / max(W_v, H_v)).
 * losses (host, num_views entries, may be NULL) receives the per-view losses. */
int msplat_hip_train_step_batch(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_views *views,
                                float ssim_weight, const msplat_hip_adam *adam, const msplat_hip_grad_stats *stats,
                                float *losses);

/* Gradients of sum_v <image_v, v_images_v> for the last render_batch on this context;
 * v_images is packed like the rendered images. Gradients are summed over the views and
 * written (not accumulated); NULL outputs are skipped. */
int msplat_hip_render_backward_batch(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_views *views,
                                     const float *v_images, const msplat_hip_grads *out);

/* Captures msplat_hip_train_step_batch into a hipGraph and replays it (1) or runs it
 * eagerly (0, the default). The graph is recaptured when the gaussian count, the view
 * count, a buffer or the intersection capacity changes; poses, resolutions within the
 * buffers, targets and Adam scalars change between replays through device memory.
 * Needs a context with its own (non-null) stream. */
int msplat_hip_set_batch_graph(msplat_hip_ctx *ctx, int enable);
/* Graph statistics: captures so far and replays so far. */
int msplat_hip_batch_graph_stats(msplat_hip_ctx *ctx, int64_t *captures, int64_t *replays);

/* Captured one-view train step (no counterpart in the reference, whose Metal backend
 * encodes a step into one command buffer, msplat_metal.mm:1698-1735). With enable = 1,
 * msplat_hip_train_step runs as a replay of a hipGraph of the whole step: no host wait
 * inside a step, one launch per step. A step whose shape differs from the current
 * graph's (gaussian count or buffers, view size, SH degree, background, loss weights,
 * Adam constants, output pointers, tuning, flags) runs launch by launch and sets the
 * intersection capacity (1.25 x its count + 4096; with initial_capacity > 0, that value
 * or the capacity an overflow has raised it to); the next step of that shape captures. Camera poses, the SH bands in use, Adam step sizes and
 * bias corrections and targets change between replays through device memory. A replay
 * over the capacity changes nothing; the library finds it at the next call, waits, runs
 * the replays that were not applied again launch by launch in order, raises the
 * capacity and recaptures, so the steps are those of the launch-by-launch path. Every
 * other call on the context (msplat_hip_sync included) first settles the replays in
 * flight; read parameters only after one of them. With MSPLAT_TUNE_CHUNKS = 0 a graph's
 * depth-chunk count is fixed at capture from the last launch-by-launch step's count.
 * Stage timing records nothing for replays. Needs a context with its own stream. */
int msplat_hip_set_step_graph(msplat_hip_ctx *ctx, int enable, uint64_t initial_capacity);
/* Captures, replays and replays run again after an overflow, so far. */
int msplat_hip_step_graph_stats(msplat_hip_ctx *ctx, int64_t *captures, int64_t *replays, int64_t *reruns);
/* Sizes of the last batched frame on this context: tile intersections and visible
 * records (views x gaussians that touch a tile). */
int msplat_hip_batch_frame_info(msplat_hip_ctx *ctx, uint64_t *num_isects, uint64_t *num_visible);

#ifdef __cplusplus
}
#endif

#endif
