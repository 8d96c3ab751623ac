// Kernel argument records and host launchers of the gfx950 kernels (internal to the
// library; the public surface is include/msplat_hip.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/msplat_hip.h"

namespace msplat {

constexpr int RS_ITEMS = 4096;  // radix-sort elements per block (256 threads x 16)

// Kernel variants of the rasterizer, picked per context (msplat_hip_set_tuning)
struct Tuning {
    int chunks = 0;       // depth chunks per tile: 0 = chosen per frame, 1 = off, 2..64 = fixed
    int bwd_ppt = 4;      // backward pixels per thread: 1, 2, 4
    int fwd_batch = 256;  // forward LDS batch depth: 64, 128, 256
    int bwd_batch = 64;   // backward LDS batch depth: 32, 64, 128
    int cull = 1;         // skip a tile's gaussians whose alpha >= 1/255 ellipse misses it
    int pack_keys = 1;    // one-view tile sort on (tile | id) keys when they fit in 32 bits
};

// Per-step values of a captured one-view train step, in device memory: uploaded before
// each replay and read by the kernels in place of their by-value copies, so that one
// graph serves every view of the same size.
struct StepParams {
    msplat_hip_camera cam;
    float step[6], bc2[6];  // Adam step sizes and bias corrections
    const float *gt;        // the target image
};

struct ProjectArgs {
    int n;
    const float *means, *scales, *quats, *dc, *rest, *opac;
    msplat_hip_camera cam;
    const StepParams *sp = nullptr;  // non-null: the camera is sp->cam
    float4 *splat_a, *splat_b;
    float *splat_blue, *depths;
    float2 *aabb;
    uint2 *rect;
    int32_t *radii;
    uint32_t *count, *block_sums, *block_vis;  // per 256-gaussian block: intersections, visible gaussians
};

constexpr uint32_t BIG_SLOTS = 32;  // records with more intersections reduce their slots per wave

struct BatchTable;

// Output of a multi-level scan: totals, and a capacity check (total > cap sets
// *overflow and makes total32 0; without overflow, total32 saturates).
struct ScanOut {
    uint64_t *total;
    uint32_t *total32;
    uint32_t cap;
    uint32_t *overflow;
    uint32_t *zero_also;  // a counter cleared by the same launch (or null)
    uint32_t *applied;    // null, or: an overflow already set stays set (this scan is over
                          // capacity too), and a scan within capacity adds one here
    uint64_t *running;    // null, or the total is added here (a running count over frames)
};

// Visible records (count > 0) in id order with their depth bits as sort keys
struct CompactArgs {
    int n, num_views, nbx;  // records r = v * n + g, nbx blocks of 256 per view
    const uint32_t *count;
    const float *depths;
    const uint32_t *block_vis;  // scanned visible counts per (view, block)
    uint32_t *dkeys;
    int32_t *dvals;
};

// Key emission over the visible records in depth order
struct KeygenArgs {
    const int32_t *list;       // visible records, depth order
    const uint32_t *nvis_ptr;  // device length of the list
    uint32_t nvis_max;         // grid bound
    const uint32_t *count;
    const uint2 *rect;
    const uint32_t *block_bases;  // scanned count sums per 256 list entries, or
    const uint32_t *chunk_sums;   // (when not null) the sums themselves, unscanned
    uint32_t *offsets;
    uint32_t *loff;         // offsets in list order (nvis_max entries)
    const uint32_t *m_ptr;  // device intersection count
    uint32_t m_max;         // grid bound of the emission
    uint32_t *keys;  // global tile id, or (tile id << gid_bits) | record id
    int32_t *vals;   // record id (unused when gid_bits > 0)
    uint32_t gid_bits;  // one view: > 0 packs the record id into the key's low bits
    uint32_t *big_list, *big_count;  // records with count > BIG_SLOTS
    uint8_t *slot_f;                 // null, or the slots' zero flags, cleared by the emission
    const uint32_t *overflow;        // null, or skip everything when set
    int tiles_x;                     // one view
    const BatchTable *tab;           // batch: record r is in view r / n
    int n;
};

struct RasterArgs {
    uint32_t width, height;
    int tiles_x, tiles_y;
    const int32_t *tile_start, *gid;
    uint32_t gid_mask;  // record id = gid[k] & gid_mask (the sort key may hold the tile id above it)
    const float4 *splat_a, *splat_b;
    const float *splat_blue;
    float bg[3];
    // forward outputs / backward inputs
    float *image, *final_T;
    int32_t *final_idx;
    // backward
    const float *v_image;
    const uint2 *rect;
    const uint32_t *offsets;
    float4 *slot_a, *slot_b;
    float *slot_c;
    uint8_t *slot_f;  // per slot: 1 = written; 0 = all nine values zero (left unwritten), and
                      // all 0 when the raster backward starts
    unsigned long long *slots_written;  // null, or a running count of written slots
    int cull;  // set by the launchers from Tuning::cull
    // depth chunks per tile (1: one workgroup per tile) and the chunked form's per-chunk
    // planes of chunks * width * height entries
    int chunks;
    float *chunk_prod;
    float4 *chunk_state;
    int32_t *chunk_last;
    // prior-guided (aux) frames (0: none): per-gaussian view depth, per-pixel background,
    // expected-depth sum out, prior-loss gradients (dL/dD, dL/dA) in, depth slots out
    int aux;
    const float *depths;
    const float *bg_img;
    float *out_depth;
    const float *v_aux;
    float *slot_d;
};

// Loss blocks: 8 rows x 118 columns; LOSS_PLANES floats of scratch per pixel (15
// horizontal SSIM moments, 9 back-convolved derivative fields, 3 SSIM terms)
constexpr int LOSS_BLOCK_ROWS = 8, LOSS_BLOCK_COLS = 118, LOSS_PLANES = 27;
inline uint32_t loss_blocks(uint32_t W, uint32_t H) {
    return ((W + LOSS_BLOCK_COLS - 1) / LOSS_BLOCK_COLS) * ((H + LOSS_BLOCK_ROWS - 1) / LOSS_BLOCK_ROWS);
}

struct LossArgs {
    uint32_t width, height;
    const float *img, *gt;
    const StepParams *sp = nullptr;  // non-null: the target is sp->gt
    float ssim_weight, inv_n;
    float *v_img;
    float *scratch;     // LOSS_PLANES planes of W x H floats
    double *row_loss;   // per row: sum, absolute sum, ulp exponent
    double *loss_sum;
    float *loss_mean;  // loss_sum / (W H), float
};

enum BackwardMode { BWD_ADAM = 0, BWD_GRADS = 1 };

struct BackwardArgs {
    int n;
    const float *means;
    float *scales, *quats, *dc, *rest, *opac, *means_rw;  // updated in place in Adam mode
    msplat_hip_camera cam;
    const StepParams *sp;      // non-null: camera, step sizes and bias corrections from here
    const uint32_t *overflow;  // non-null and set: the frame was over capacity, Adam does nothing
    // prior-guided frames: the intersections' depth gradient slots (null: none) and the
    // big records' sums of them; their sum v_depth moves the mean along the view axis
    const float *slot_d;
    float *big_d;
    const int32_t *radii;
    const uint32_t *count, *offsets;
    const float4 *splat_a, *splat_b;
    const float4 *slot_a, *slot_b;
    const float *slot_c;
    const uint8_t *slot_f;  // slots with a zero flag are zero and not read
    // Adam mode
    float *m1[6], *m2[6];
    float step[6], bc2[6];
    float beta1, beta2, eps;
    float *vis, *gn, *ms;
    float inv_max_dim;
    // gradient mode
    float *g_means, *g_scales, *g_quats, *g_dc, *g_rest, *g_opac;
    // optional per-gaussian reduced gradients (test hook)
    float *v_xy, *v_conic, *v_rgb, *v_opac;
    // per-gaussian view direction (w = visible) and colour gradient, for the SH bands
    float4 *sh_dir, *sh_vrgb;
    // slot sums of the gaussians with many intersections, reduced one wave each
    const uint32_t *big_list, *big_count;
    float4 *big_a, *big_b;
    float *big_c;
    // the visible gaussians (depth order) and their device-side count, for the SH Adam
    const int32_t *vis_list;
    const uint32_t *vis_count;
};

struct DensifyArgs {
    int n, n_new, n_split;
    float grad_thresh, size_thresh, screen_thresh, cull_alpha, cull_scale, cull_screen, half_max_dim, log_shrink;
    int check_screen, check_huge;
    uint32_t seed;
    const float *gn, *vc, *ms;
    float *bufs[18];
    int width[6];
    uint8_t *split, *dup, *keep;
    uint32_t *split_bs, *dup_bs, *keep_bs;
    int32_t *dst;
};
void launch_densify_classify(const DensifyArgs &a, hipStream_t s);
void launch_densify_append(const DensifyArgs &a, hipStream_t s);
void launch_densify_cull(const DensifyArgs &a, hipStream_t s);
void launch_densify_rank(const DensifyArgs &a, hipStream_t s);
void launch_compact_scatter(const float *src, float *scratch, const int32_t *dst, int n_new, int width, hipStream_t s);
void launch_opacity_reset(float *o, float *m1, float *m2, int n, float reset, hipStream_t s);
void launch_pack_rgba8(const float *img, uint32_t n, uint8_t *out, hipStream_t s);
void launch_metrics(const float *a, const float *b, int h, int w, float *mom, float *ssim, double *rows, double *out,
                    hipStream_t s);

struct KnnScratch {
    float *bounds, *box, *sup;
    uint64_t *codes, *codes_tmp;
    int32_t *order, *order_tmp;
    float4 *sorted;
    uint32_t *hist, *digit_total, *n_dev;
};
size_t knn_scratch_bytes(uint32_t n);
void knn_scratch_carve(void *base, uint32_t n, KnnScratch &w);
int knn3_launch(const float *pts, uint32_t n, float *mean_dist, KnnScratch &w, hipStream_t s);
void launch_init_gaussians(const float *md, const uint8_t *rgb, uint32_t n, uint32_t seed, float logit, float *scales,
                           float *quats, float *dc, float *opac, hipStream_t s);

// ── prior-guided extras (PriorStep, core/metal/bindings.h:43-77) ────────────────
// The learned sky's equirect frame and texture and the camera it is seen from
struct SkyArgs {
    float up[3], e1[3], e2[3];
    float fx, fy, cx, cy;
    float vm[12];  // world-to-camera rows 0-2
    const float *tex;  // (h, w, 3), or null: the constant background
    uint32_t w, h;
};

// The prior losses of one frame (prior_loss, cpu:809-884)
struct PriorLossArgs {
    uint32_t width, height;
    const float *final_T, *out_depth;
    float *v_photo;  // the photometric gradient, zeroed where the prior does not keep the pixel
    float *v_aux;    // out: dL/dD, dL/dA per pixel
    double *rows;    // out: per row the depth, sky and fill terms
    const float *prior_depth;  // (ph, pw), or null: no prior image (every pixel kept, no sky)
    const uint8_t *prior_aux;  // (ph, pw, 4): confidence, sky, keep, -
    uint32_t pw, ph;
    int has_depth, has_sky;
    float depth_weight, sky_weight, fill_weight, huber, min_alpha;
    int mask_photometric;
};

void launch_sky_background(const SkyArgs &k, uint32_t W, uint32_t H, const float bg[3], float *bg_img, hipStream_t s);
void launch_exposure_apply(const float *img, const float *P, size_t n, float *out, hipStream_t s);
void launch_prior_loss(const PriorLossArgs &a, float *loss_terms, hipStream_t s);
// v_img = P^T v_adj per pixel, the 12 gradient sums (double partials in `part`, one row of
// 12 per block), then one Adam step on the camera's row (`row` = params + 12 cam)
void launch_exposure_backward_adam(const float *img, const float *v_adj, size_t n, float *v_img, double *part,
                                   float *params, float *m1, float *m2, float *grad, int cam, float step, float bc2,
                                   float reg, float b1, float b2, float eps, hipStream_t s);
size_t exposure_partials(size_t n);
// the sky texture's gradient from T v_img at each pixel's taps (fixed-point sums in acc,
// w x h x 3 words, order-independent), written to grad, then Adam
void launch_sky_backward_adam(const SkyArgs &k, uint32_t W, uint32_t H, const float *final_T, const float *v_img,
                              unsigned long long *acc, float *grad, float *tex, float *m1, float *m2, float step,
                              float bc2, float b1, float b2, float eps, hipStream_t s);
void launch_scale_cap(float *scales, int n, float log_max, hipStream_t s);
void launch_finalize_depth(const float *depth_num, const float *final_T, uint32_t n, float inv_scale,
                           float *depth_out, float *alpha_out, hipStream_t s);

// ── batched views ───────────────────────────────────────────────────────────
// Several views of one scene go through one projection, one depth sort of the visible
// records (then stably by view), one sort of the intersections by local tile id (each
// tile's range then holds the views in order, each in depth order: range index
// tile * num_views + view), one persistent rasterizer over (view, tile) items, one loss
// launch and one per-gaussian backward that sums over the views. Everything that
// changes from step to step (poses, resolutions, targets, Adam scalars) is read from a
// BatchTable in device memory, so a captured graph replays with new contents.
// Record r = v * n + g holds gaussian g's projection in view v.
constexpr int MAX_VIEWS = MSPLAT_MAX_BATCH_VIEWS;

struct BatchView {
    msplat_hip_camera cam;
    const float *gt;  // (H, W, 3) target (training)
    float bg[3];
    float inv_n;        // loss normalisation, 1 / (3 W H)
    float inv_max_dim;  // screen-size statistic scale
    uint32_t pix_base;  // first pixel of the view in the packed per-pixel buffers
};

struct BatchTable {
    int32_t num_views;
    uint32_t total_tiles, total_pixels, total_rows, total_loss_blocks;
    float ssim_weight;
    float step[6], bc2[6];
    float beta1, beta2, eps;
    uint32_t rest_active;  // SH rest coefficients per gaussian in the bands in use (max over the views)
    uint32_t max_view_tiles;  // tiles of the largest view: ranges run over max_view_tiles * num_views
    // prefix offsets over the views (num_views + 1 entries each)
    uint32_t tile_base[MAX_VIEWS + 1], row_base[MAX_VIEWS + 1], loss_base[MAX_VIEWS + 1];
    BatchView views[MAX_VIEWS];
};

// Device scalars of a batched frame
struct BatchScalars {
    uint64_t total;      // intersections
    uint64_t nvis64;     // visible records
    uint32_t total32;    // intersections, 0 when they exceed the capacity
    uint32_t overflow;   // 1 when the intersections exceed the capacity: the step is void
    uint32_t big_count;  // records with more than BIG_SLOTS intersections
    uint32_t nvis;       // visible records
    uint32_t queue[2][8];  // persistent rasterizer queue heads (forward, backward) per XCD chunk
    uint32_t total32_b, overflow_b;  // the same totals from the depth-order scan
    double loss_sum[MAX_VIEWS];
    float loss_mean[MAX_VIEWS];
};

struct BatchProjectArgs {
    const BatchTable *tab;
    int n, num_views, nbx;  // nbx = ceil(n / 256) blocks per view
    const float *means, *scales, *quats, *dc, *rest, *opac;
    float4 *splat_a, *splat_b;
    float *splat_blue, *depths;
    uint2 *rect;
    int32_t *radii;
    uint32_t *count, *block_sums, *block_vis;  // [v * nbx + b]
    int sh_u;          // SH floats staged per gaussian: 3 sh_bases(largest degree in use of the views);
                       // 0 (read in place) at degree 4 or with only the DC in use
    size_t sh_stride;  // rest floats per gaussian
};

struct BatchLossArgs {
    const BatchTable *tab;
    int num_views;
    uint32_t max_loss_blocks, max_rows;
    const float *img;
    float *v_img;
    float *scratch;    // LOSS_PLANES floats per pixel, view v's planes at LOSS_PLANES pix_base
    double *row_loss;  // 3 doubles per row
    BatchScalars *sc;
};

struct BatchBackwardArgs {
    const BatchTable *tab;
    int n, num_views;
    const float *means, *scales, *quats;
    const int32_t *radii;
    const uint32_t *count, *offsets;
    const float4 *splat_a, *splat_b;
    const float4 *slot_a, *slot_b;
    const float *slot_c;
    const uint8_t *slot_f;  // slots with a zero flag are zero and not read
    const float4 *big_a, *big_b;  // slot sums of records with count > BIG_SLOTS
    const float *big_c;
    const BatchScalars *sc;
    uint32_t rest_stride;  // stored (bases - 1) * 3
    // gradients summed over the views, and whether the gaussian is visible in any view
    float *g_means, *g_scales, *g_quats, *g_dc, *g_rest, *g_opac;
    uint8_t *vis_any;
    // densification statistics (null: not updated)
    float *vis, *gn, *ms;
};

void launch_batch_project(const BatchProjectArgs &a, hipStream_t s);
struct BatchRasterArgs {
    const BatchTable *tab;
    int num_views;
    uint32_t *queue;
    const int32_t *tile_start, *gid;
    const float4 *splat_a, *splat_b;
    const float *splat_blue;
    const uint2 *rect;
    const uint32_t *offsets;
    float *image, *final_T;
    int32_t *final_idx;
    const float *v_image;
    float4 *slot_a, *slot_b;
    float *slot_c;
    uint8_t *slot_f;  // per slot: 1 = written; 0 = all nine values zero (left unwritten), and
                      // all 0 when the raster backward starts
    int cull;  // set by the launchers from Tuning::cull
};
int batch_raster_grid(bool backward, const Tuning &t);  // persistent workgroups that fill the device
void launch_batch_raster_forward(const BatchRasterArgs &a, const Tuning &t, hipStream_t s);
void launch_batch_raster_backward(const BatchRasterArgs &a, const Tuning &t, hipStream_t s);
void launch_batch_loss(const BatchLossArgs &a, hipStream_t s);
void launch_big_reduce_records(const uint32_t *big_list, const uint32_t *big_count, const uint32_t *offsets,
                               const uint32_t *count, const float4 *slot_a, const float4 *slot_b, const float *slot_c,
                               const uint8_t *slot_f, float4 *big_a, float4 *big_b, float *big_c, hipStream_t s);
void launch_batch_backward(const BatchBackwardArgs &a, hipStream_t s);
// Adam over one parameter group from summed gradients; `vis` (per gaussian, null =
// dense) restricts it to gaussians visible in some view, and for the SH rest group to
// the coefficients of the bands in use; void when overflow is set.
void launch_batch_adam(float *p, float *m1, float *m2, const float *grad, size_t count, int width,
                       const uint8_t *vis, const BatchTable *tab, int group, const uint32_t *overflow, hipStream_t s);

void launch_project(const ProjectArgs &a, hipStream_t s);
void launch_rect_count(int n, const float4 *splat_a, const float2 *aabb, const int32_t *radii, int tx, int ty,
                       uint2 *rect, uint32_t *count, uint32_t *block_sums, uint32_t *block_vis, hipStream_t s);
void launch_scan_blocks(uint32_t *sums, int nblocks, uint64_t *total, uint32_t *total32, hipStream_t s,
                        uint32_t *zero_also = nullptr);
size_t scan_partial_words(uint32_t L);
void launch_scan_u32(uint32_t *a, uint32_t L, uint32_t *partial, const ScanOut &o, hipStream_t s);
// two scans of length L (one launch when L is small)
void launch_scan_u32_pair(uint32_t *a0, const ScanOut &o0, uint32_t *a1, const ScanOut &o1, uint32_t L,
                          uint32_t *partial0, uint32_t *partial1, hipStream_t s);
void launch_compact(const CompactArgs &a, hipStream_t s);
// also writes each record's count in list order to lcount (the offsets kernel's loff)
void launch_order_sums(const int32_t *list, const uint32_t *nvis_dev, uint32_t nvis_max, const uint32_t *count,
                       uint32_t *sums, uint32_t *lcount, hipStream_t s);
void launch_keygen_offsets(const KeygenArgs &a, hipStream_t s);  // offsets, loff, big list
void launch_keygen_emit(const KeygenArgs &a, bool batch, hipStream_t s);  // the keys
void launch_keygen(const KeygenArgs &a, bool batch, hipStream_t s);
// One stable pass over the digit (key >> shift) & mask (mask at most 0xFF)
void launch_radix_pass(const uint64_t *kin, const int32_t *vin, uint64_t *kout, int32_t *vout, const uint32_t *n_dev,
                       uint32_t n_max, int shift, uint32_t *hist, uint32_t *digit_total, hipStream_t s,
                       uint32_t mask = 0xFF);
void launch_radix_pass(const uint32_t *kin, const int32_t *vin, uint32_t *kout, int32_t *vout, const uint32_t *n_dev,
                       uint32_t n_max, int shift, uint32_t *hist, uint32_t *digit_total, hipStream_t s,
                       uint32_t mask = 0xFF);
size_t radix_hist_words(uint32_t n_max);
void launch_tile_ranges(const uint32_t *keys, const uint32_t *n_dev, uint32_t n_max, const uint32_t *ntiles_dev,
                        int ntiles, int32_t *tile_start, hipStream_t s, int shift = 0);
// tile ids (key >> shift) and record ids (key & (2^shift - 1)) of packed sorted keys
void launch_unpack_keys(const uint32_t *packed, uint32_t n, int shift, uint32_t *tiles, int32_t *gid, hipStream_t s);
void launch_keys64(const uint32_t *tiles, const int32_t *gid, const float *depths, uint32_t n, uint64_t *out,
                   hipStream_t s);
// batch: view ids of the depth-sorted records (record / n) as sort keys
void launch_view_keys(const int32_t *recs, const uint32_t *n_dev, uint32_t n_max, int n, uint32_t *keys,
                      hipStream_t s);
// batch: ranges over (local tile, view) pairs, index tile * num_views + view
void launch_tile_ranges_views(const uint32_t *keys, const int32_t *recs, const uint32_t *n_dev, uint32_t n_max, int n,
                              const BatchTable *tab, int32_t *range_start, hipStream_t s);
int auto_chunks(int nt, int64_t m);
void launch_raster_forward(const RasterArgs &a, const Tuning &t, hipStream_t s);
void launch_raster_backward(const RasterArgs &a, const Tuning &t, hipStream_t s);
void launch_loss(const LossArgs &a, hipStream_t s);
void launch_backward(const BackwardArgs &a, BackwardMode mode, hipStream_t s);

}  // namespace msplat
