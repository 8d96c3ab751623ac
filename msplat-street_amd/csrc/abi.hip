// C ABI of the HIP backend (include/msplat_hip.h): contexts, device memory, and the
// render / train / backward pipelines over the gfx950 kernels.
//
// A context owns one stream and a grow-only workspace: per-gaussian projection records,
// per-intersection keys/ids/gradient slots, per-tile ranges and per-pixel buffers. The
// intersection count is read back once per view (after the per-block count scan) to
// size the sort; everything else stays on the device.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "context.h"

using namespace msplat;

namespace {
thread_local std::string g_last_error;
}  // namespace

int msplat::set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

namespace {

// projection (per-block intersection and visible counts)
int do_project(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_camera *cam,
               const StepParams *sp = nullptr) {
    ctx->frame_aux = false;  // a prior-guided entry point sets it after the projection
    const int n = g->num_points;
    const size_t N = (size_t)(n > 0 ? n : 1);
    const int nblocks = (n + 255) / 256;
    CHECKED(ctx->grow(ctx->splat_a, 16 * N));
    CHECKED(ctx->grow(ctx->splat_b, 16 * N));
    CHECKED(ctx->grow(ctx->splat_blue, 4 * N));
    CHECKED(ctx->grow(ctx->depths, 4 * N));
    CHECKED(ctx->grow(ctx->aabb, 8 * N));
    CHECKED(ctx->grow(ctx->rect, 8 * N));
    CHECKED(ctx->grow(ctx->radii, 4 * N));
    CHECKED(ctx->grow(ctx->count, 4 * N));
    CHECKED(ctx->grow(ctx->offsets, 4 * N));
    CHECKED(ctx->grow(ctx->block_sums, 4 * (size_t)(nblocks + 1)));
    CHECKED(ctx->grow(ctx->block_vis, 4 * (size_t)(nblocks + 1)));
    if (!ctx->scalars.p) {
        CHECKED(ctx->grow(ctx->scalars, 128));
        zero_u32<<<1, 256, 0, ctx->stream>>>((uint32_t *)ctx->scalars.p, 32);
    }
    CHECKED(ctx->grow(ctx->big, 40 * N + 64));  // list (4) + slot sums (16 + 16 + 4) + alignment
    ctx->n = n;
    ctx->width = cam->width;
    ctx->height = cam->height;
    ctx->tiles_x = cam->tiles_x;
    ctx->tiles_y = cam->tiles_y;
    ctx->ntiles = cam->tiles_x * cam->tiles_y;
    ProjectArgs a;
    a.n = n;
    a.means = g->means;
    a.scales = g->scales;
    a.quats = g->quats;
    a.dc = g->features_dc;
    a.rest = g->features_rest;
    a.opac = g->opacities;
    a.cam = *cam;
    a.sp = sp;
    a.splat_a = (float4 *)ctx->splat_a.p;
    a.splat_b = (float4 *)ctx->splat_b.p;
    a.splat_blue = (float *)ctx->splat_blue.p;
    a.depths = (float *)ctx->depths.p;
    a.aabb = (float2 *)ctx->aabb.p;
    a.rect = (uint2 *)ctx->rect.p;
    a.radii = (int32_t *)ctx->radii.p;
    a.count = (uint32_t *)ctx->count.p;
    a.block_sums = (uint32_t *)ctx->block_sums.p;
    a.block_vis = (uint32_t *)ctx->block_vis.p;
    ctx->begin(0);
    launch_project(a, ctx->stream);
    HIPCHECK(hipGetLastError());
    ctx->frame_stage = msplat_hip_ctx::FRAME_PROJECTED;
    return MSPLAT_OK;
}

// Count scans and their read-back (M intersections, V visible gaussians), then the
// visible gaussians sorted by depth, keys emitted in that order, the stable tile sort
// and the tile ranges. The host waits for the counts only after the depth sort and
// the offsets are enqueued (those run on device-side counts against the gaussian
// capacity), so the device sorts while the host reads M and sizes the key buffers.
// With an intersection capacity `cap` (a captured step) nothing is read back: every
// launch after the counts is sized for the capacity and reads the device-side count,
// and a frame over capacity sets the overflow word and leaves an empty tile list.
int do_bin(msplat_hip_ctx *ctx, uint32_t cap = 0) {
    if (ctx->frame_stage < msplat_hip_ctx::FRAME_PROJECTED)
        return set_error(MSPLAT_ERR_ARG, "binning needs a projected frame on this context");
    ctx->frame_stage = msplat_hip_ctx::FRAME_PROJECTED;  // the key buffers are rewritten below
    const int n = ctx->n;
    const uint32_t nblocks = (uint32_t)((n + 255) / 256);
    if (!ctx->h_total) HIPCHECK(hipHostMalloc((void **)&ctx->h_total, 128, hipHostMallocDefault));
    if (!ctx->counts_ready) HIPCHECK(hipEventCreateWithFlags(&ctx->counts_ready, hipEventDisableTiming));
    const size_t Nc = n > 0 ? (size_t)n : 1;
    const uint32_t vblocks = (uint32_t)((Nc + 255) / 256);  // depth-ordered list blocks, at most
    CHECKED(ctx->grow(ctx->scan_part, 4 * 3 * scan_partial_words(std::max(nblocks, vblocks)) + 64));
    CHECKED(ctx->grow(ctx->hist, 4 * radix_hist_words((uint32_t)std::max<size_t>(Nc, ctx->M > 0 ? ctx->M : 1))));
    CHECKED(ctx->grow(ctx->digit_total, 4 * 256));
    CHECKED(ctx->grow(ctx->dk0, 4 * Nc));
    CHECKED(ctx->grow(ctx->dk1, 4 * Nc));
    CHECKED(ctx->grow(ctx->dv0, 4 * Nc));
    CHECKED(ctx->grow(ctx->dv1, 4 * Nc));
    CHECKED(ctx->grow(ctx->ord_sums, 4 * (size_t)(vblocks + 1)));
    CHECKED(ctx->grow(ctx->loff, 4 * Nc));
    CHECKED(ctx->grow(ctx->tile_start, 4 * (size_t)(ctx->ntiles + 1)));
    uint32_t *part = (uint32_t *)ctx->scan_part.p;
    const size_t pw = scan_partial_words(std::max(nblocks, vblocks));
    hipStream_t s = ctx->stream;
    const ScanOut isects = cap ? ScanOut{ctx->d_total64(), ctx->d_total32(), cap, ctx->d_overflow(), ctx->d_big_count(),
                                         ctx->d_applied(), ctx->d_isects_running()}
                               : ScanOut{ctx->d_total64(), ctx->d_total32(), 0xFFFFFFFFu, nullptr, ctx->d_big_count(),
                                         nullptr, ctx->d_isects_running()};
    launch_scan_u32_pair((uint32_t *)ctx->block_sums.p, isects, (uint32_t *)ctx->block_vis.p,
                         ScanOut{ctx->d_vis64(), ctx->d_vis32(), 0xFFFFFFFFu, nullptr, nullptr, nullptr,
                                 ctx->d_visible_running()},
                         nblocks, part, part + pw, s);
    ctx->end();
    if (!cap) {
        HIPCHECK(hipMemcpyAsync(ctx->h_total, ctx->scalars.p, 64, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipEventRecord(ctx->counts_ready, s));
    }
    ctx->begin(1);
    // visible gaussians in id order, stably sorted by depth bits
    CompactArgs c;
    c.n = n;
    c.num_views = 1;
    c.nbx = (int)nblocks;
    c.count = (const uint32_t *)ctx->count.p;
    c.depths = (const float *)ctx->depths.p;
    c.block_vis = (const uint32_t *)ctx->block_vis.p;
    c.dkeys = (uint32_t *)ctx->dk0.p;
    c.dvals = (int32_t *)ctx->dv0.p;
    launch_compact(c, s);
    uint32_t *dki = (uint32_t *)ctx->dk0.p, *dko = (uint32_t *)ctx->dk1.p;
    int32_t *dvi = (int32_t *)ctx->dv0.p, *dvo = (int32_t *)ctx->dv1.p;
    for (int shift = 0; shift < 32; shift += 8) {
        launch_radix_pass(dki, dvi, dko, dvo, ctx->d_vis32(), (uint32_t)Nc, shift, (uint32_t *)ctx->hist.p,
                          (uint32_t *)ctx->digit_total.p, s);
        std::swap(dki, dko);
        std::swap(dvi, dvo);
    }
    ctx->vis_sorted = dvi;
    // intersection offsets in depth order
    launch_order_sums(dvi, ctx->d_vis32(), (uint32_t)Nc, (const uint32_t *)ctx->count.p, (uint32_t *)ctx->ord_sums.p,
                      (uint32_t *)ctx->loff.p, s);
    KeygenArgs k;
    memset(&k, 0, sizeof k);
    k.list = dvi;
    k.nvis_ptr = ctx->d_vis32();
    k.nvis_max = (uint32_t)Nc;
    k.count = (const uint32_t *)ctx->count.p;
    k.rect = (const uint2 *)ctx->rect.p;
    // up to 4096 chunks the key generation adds the preceding chunk sums itself (a few
    // loads per thread); past that, one scan launch is cheaper than the quadratic reads
    if (vblocks <= 4096) {
        k.chunk_sums = (const uint32_t *)ctx->ord_sums.p;
    } else {
        launch_scan_u32((uint32_t *)ctx->ord_sums.p, vblocks, part + 2 * pw,
                        ScanOut{nullptr, nullptr, 0xFFFFFFFFu, nullptr, nullptr}, s);
        k.block_bases = (const uint32_t *)ctx->ord_sums.p;
    }
    k.offsets = (uint32_t *)ctx->offsets.p;
    k.loff = (uint32_t *)ctx->loff.p;
    k.m_ptr = ctx->d_total32();
    k.big_list = ctx->big_list();
    k.big_count = ctx->d_big_count();
    k.tiles_x = ctx->tiles_x;
    k.overflow = cap ? ctx->d_overflow() : nullptr;
    launch_keygen_offsets(k, s);
    HIPCHECK(hipGetLastError());
    // the counts (or the capacity): size the key buffers, then keys (tile ids), stable
    // tile sort, ranges
    uint64_t M = cap;
    if (!cap) {
        HIPCHECK(hipEventSynchronize(ctx->counts_ready));
        ctx->collect();
        M = ctx->h_total[0];
        if (M > 0x7FFFFFFFull)
            return set_error(MSPLAT_ERR_RANGE, "more than 2^31 tile intersections (%llu); lower the resolution",
                             (unsigned long long)M);
        ctx->M = (int64_t)M;
        ctx->V = (int64_t)ctx->h_total[5];
    }
    const size_t Mc = M > 0 ? M : 1;
    CHECKED(ctx->grow(ctx->keys0, 4 * Mc));
    CHECKED(ctx->grow(ctx->keys1, 4 * Mc));
    CHECKED(ctx->grow(ctx->vals0, 4 * Mc));
    CHECKED(ctx->grow(ctx->vals1, 4 * Mc));
    CHECKED(ctx->grow(ctx->hist, 4 * radix_hist_words((uint32_t)std::max(Mc, Nc))));
    CHECKED(ctx->grow(ctx->slot_f, (Mc + 3) & ~(size_t)3));
    // (tile << gbits) | id fits in 32 bits: the sort moves the keys alone
    const int bits = tile_bits(ctx->ntiles);
    int gbits = 1;
    while (gbits < 31 && (1u << gbits) < (uint32_t)Nc) gbits++;
    const bool packed = ctx->tuning.pack_keys && gbits + bits <= 32;
    k.m_max = (uint32_t)M;
    k.keys = (uint32_t *)ctx->keys0.p;
    k.vals = packed ? nullptr : (int32_t *)ctx->vals0.p;
    k.gid_bits = packed ? (uint32_t)gbits : 0u;
    k.slot_f = (uint8_t *)ctx->slot_f.p;
    launch_keygen_emit(k, false, s);
    ctx->slot_f_clean = true;
    uint32_t *kin = (uint32_t *)ctx->keys0.p, *kout = (uint32_t *)ctx->keys1.p;
    int32_t *vin = packed ? nullptr : (int32_t *)ctx->vals0.p, *vout = packed ? nullptr : (int32_t *)ctx->vals1.p;
    const int shift0 = packed ? gbits : 0;
    // The first pass takes the bits left over by whole bytes (the last pass takes 8), so
    // that it has fewer, longer digit runs: its input is in depth order, the least
    // coalesced scatter of the sort.
    for (int done = 0; done < bits;) {
        const int width = done == 0 && bits % 8 ? bits % 8 : 8;
        launch_radix_pass(kin, vin, kout, vout, ctx->d_total32(), (uint32_t)M, shift0 + done,
                          (uint32_t *)ctx->hist.p, (uint32_t *)ctx->digit_total.p, s, (1u << width) - 1u);
        std::swap(kin, kout);
        std::swap(vin, vout);
        done += width;
    }
    ctx->keys_sorted = kin;
    ctx->gid_sorted = vin;
    ctx->sort_gid_bits = shift0;
    launch_tile_ranges(kin, ctx->d_total32(), (uint32_t)M, nullptr, ctx->ntiles, (int32_t *)ctx->tile_start.p, s,
                       shift0);
    ctx->end();
    HIPCHECK(hipGetLastError());
    ctx->frame_stage = msplat_hip_ctx::FRAME_BINNED;
    return MSPLAT_OK;
}

int ensure_pixels(msplat_hip_ctx *ctx) {
    const size_t P = (size_t)ctx->width * ctx->height;
    CHECKED(ctx->grow(ctx->image, 12 * P));
    CHECKED(ctx->grow(ctx->final_T, 4 * P));
    CHECKED(ctx->grow(ctx->final_idx, 4 * P));
    CHECKED(ctx->grow(ctx->v_image, 12 * P));
    CHECKED(ctx->grow(ctx->row_loss, 24 * (size_t)ctx->height));
    return MSPLAT_OK;
}

int ensure_loss_scratch(msplat_hip_ctx *ctx) {
    return ctx->grow(ctx->loss_scratch, 4 * (size_t)LOSS_PLANES * ctx->width * ctx->height);
}

RasterArgs raster_args(msplat_hip_ctx *ctx) {
    RasterArgs r;
    r.width = ctx->width;
    r.height = ctx->height;
    r.tiles_x = ctx->tiles_x;
    r.tiles_y = ctx->tiles_y;
    r.tile_start = (const int32_t *)ctx->tile_start.p;
    r.gid = ctx->sort_gid_bits ? (const int32_t *)ctx->keys_sorted : ctx->gid_sorted;
    r.gid_mask = ctx->sort_gid_bits ? (1u << ctx->sort_gid_bits) - 1u : 0xFFFFFFFFu;
    r.splat_a = (const float4 *)ctx->splat_a.p;
    r.splat_b = (const float4 *)ctx->splat_b.p;
    r.splat_blue = (const float *)ctx->splat_blue.p;
    for (int i = 0; i < 3; i++) r.bg[i] = ctx->bg[i];
    r.image = (float *)ctx->image.p;
    r.final_T = (float *)ctx->final_T.p;
    r.final_idx = (int32_t *)ctx->final_idx.p;
    r.v_image = (const float *)ctx->v_image.p;
    r.rect = (const uint2 *)ctx->rect.p;
    r.offsets = (const uint32_t *)ctx->offsets.p;
    r.slot_a = (float4 *)ctx->slot_a.p;
    r.slot_b = (float4 *)ctx->slot_b.p;
    r.slot_c = (float *)ctx->slot_c.p;
    r.slot_f = (uint8_t *)ctx->slot_f.p;
    r.slots_written = nullptr;
    r.chunks = ctx->frame_chunks;
    r.chunk_prod = (float *)ctx->chunk_prod.p;
    r.chunk_state = (float4 *)ctx->chunk_state.p;
    r.chunk_last = (int32_t *)ctx->chunk_last.p;
    r.aux = ctx->frame_aux ? 1 : 0;
    r.depths = ctx->frame_aux ? (const float *)ctx->depths.p : nullptr;
    r.bg_img = ctx->frame_aux ? (const float *)ctx->bg_img.p : nullptr;
    r.out_depth = ctx->frame_aux ? (float *)ctx->out_depth.p : nullptr;
    r.v_aux = ctx->frame_aux ? (const float *)ctx->v_aux.p : nullptr;
    r.slot_d = ctx->frame_aux ? (float *)ctx->slot_d.p : nullptr;
    return r;
}

int do_raster_forward(msplat_hip_ctx *ctx, const float bg[3]) {
    if (ctx->frame_stage < msplat_hip_ctx::FRAME_BINNED)
        return set_error(MSPLAT_ERR_ARG, "forward rasterization needs a binned frame on this context");
    for (int i = 0; i < 3; i++) ctx->bg[i] = bg ? bg[i] : 0.f;
    CHECKED(ensure_pixels(ctx));
    const int tc = ctx->tuning.chunks;
    ctx->frame_chunks = ctx->M <= 0 || ctx->frame_aux ? 1 : tc == 0 ? auto_chunks(ctx->ntiles, ctx->M) : tc;
    if (ctx->frame_chunks > 1) {
        const size_t planes = (size_t)ctx->frame_chunks * ctx->width * ctx->height;
        CHECKED(ctx->grow(ctx->chunk_prod, 4 * planes));
        CHECKED(ctx->grow(ctx->chunk_state, 16 * planes));
        CHECKED(ctx->grow(ctx->chunk_last, 4 * planes));
    }
    ctx->begin(2);
    launch_raster_forward(raster_args(ctx), ctx->tuning, ctx->stream);
    ctx->end();
    HIPCHECK(hipGetLastError());
    ctx->frame_stage = msplat_hip_ctx::FRAME_RASTERED;
    return MSPLAT_OK;
}

// The photometric loss of the frame's image (or of `img`, with its gradient into `v_img`)
int do_loss(msplat_hip_ctx *ctx, const float *gt, float ssim_weight, float inv_n, const StepParams *sp = nullptr,
            const float *img = nullptr, float *v_img = nullptr) {
    if (ctx->frame_stage < msplat_hip_ctx::FRAME_RASTERED)
        return set_error(MSPLAT_ERR_ARG, "the loss needs a rendered frame on this context");
    LossArgs l;
    l.width = ctx->width;
    l.height = ctx->height;
    l.img = img ? img : (const float *)ctx->image.p;
    l.gt = gt;
    l.sp = sp;
    l.ssim_weight = ssim_weight;
    l.inv_n = inv_n;
    l.v_img = v_img ? v_img : (float *)ctx->v_image.p;
    CHECKED(ensure_loss_scratch(ctx));
    l.scratch = (float *)ctx->loss_scratch.p;
    l.row_loss = (double *)ctx->row_loss.p;
    l.loss_sum = ctx->d_loss_sum();
    l.loss_mean = ctx->d_loss();
    ctx->begin(3);
    launch_loss(l, ctx->stream);
    ctx->end();
    HIPCHECK(hipGetLastError());
    return MSPLAT_OK;
}

int do_raster_backward(msplat_hip_ctx *ctx) {
    if (ctx->frame_stage < msplat_hip_ctx::FRAME_RASTERED)
        return set_error(MSPLAT_ERR_ARG, "the backward needs a rendered frame on this context");
    const size_t Mc = ctx->capturing ? (size_t)ctx->step_cap : ctx->M > 0 ? (size_t)ctx->M : 1;
    CHECKED(ctx->grow(ctx->slot_a, 16 * Mc));
    CHECKED(ctx->grow(ctx->slot_b, 16 * Mc));
    CHECKED(ctx->grow(ctx->slot_c, 4 * Mc));
    CHECKED(ctx->grow(ctx->slot_f, (Mc + 3) & ~(size_t)3));
    if (ctx->frame_aux) CHECKED(ctx->grow(ctx->slot_d, 4 * Mc));
    // the zero flags of the frame's slots: cleared by its key emission, or here when an
    // earlier backward of the same frame has set some
    if (!ctx->slot_f_clean) {
        const size_t words = (Mc + 3) / 4;
        zero_u32<<<(unsigned)std::min<size_t>((words + 255) / 256, 2048), 256, 0, ctx->stream>>>(
            (uint32_t *)ctx->slot_f.p, words);
    }
    RasterArgs r = raster_args(ctx);
    r.slots_written = ctx->d_slots_written();
    ctx->begin(4);
    launch_raster_backward(r, ctx->tuning, ctx->stream);
    ctx->end();
    ctx->slot_f_clean = false;
    HIPCHECK(hipGetLastError());
    return MSPLAT_OK;
}

int backward_args(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_camera *cam, BackwardArgs &b) {
    memset(&b, 0, sizeof b);
    b.n = g->num_points;
    b.means = g->means;
    b.means_rw = g->means;
    b.scales = g->scales;
    b.quats = g->quats;
    b.dc = g->features_dc;
    b.rest = g->features_rest;
    b.opac = g->opacities;
    b.cam = *cam;
    b.radii = (const int32_t *)ctx->radii.p;
    b.count = (const uint32_t *)ctx->count.p;
    b.offsets = (const uint32_t *)ctx->offsets.p;
    b.splat_a = (const float4 *)ctx->splat_a.p;
    b.splat_b = (const float4 *)ctx->splat_b.p;
    b.slot_a = (const float4 *)ctx->slot_a.p;
    b.slot_b = (const float4 *)ctx->slot_b.p;
    b.slot_c = (const float *)ctx->slot_c.p;
    b.slot_f = (const uint8_t *)ctx->slot_f.p;
    {
        const size_t N = (size_t)(b.n > 0 ? b.n : 1);
        CHECKED(ctx->grow(ctx->sh_scratch, 32 * N));
        b.sh_dir = (float4 *)ctx->sh_scratch.p;
        b.sh_vrgb = b.sh_dir + N;
        b.vis_list = ctx->vis_sorted;
        b.vis_count = ctx->d_vis32();
        b.big_list = ctx->big_list();
        b.big_count = ctx->d_big_count();
        float4 *ba = (float4 *)(((uintptr_t)ctx->big_a() + 15) & ~(uintptr_t)15);
        b.big_a = ba;
        b.big_b = ba + N;
        b.big_c = (float *)(ba + 2 * N);
        if (ctx->frame_aux) {
            CHECKED(ctx->grow(ctx->big_d, 4 * N));
            b.slot_d = (const float *)ctx->slot_d.p;
            b.big_d = (float *)ctx->big_d.p;
        }
    }
    if (ctx->flags & MSPLAT_FRAME_KEEP_GRADS) {
        const size_t N = (size_t)(b.n > 0 ? b.n : 1);
        CHECKED(ctx->grow(ctx->v_xy, 8 * N));
        CHECKED(ctx->grow(ctx->v_conic, 12 * N));
        CHECKED(ctx->grow(ctx->v_rgb, 12 * N));
        CHECKED(ctx->grow(ctx->v_opac, 4 * N));
        b.v_xy = (float *)ctx->v_xy.p;
        b.v_conic = (float *)ctx->v_conic.p;
        b.v_rgb = (float *)ctx->v_rgb.p;
        b.v_opac = (float *)ctx->v_opac.p;
    }
    return MSPLAT_OK;
}

// The step's projection, SH and Adam backward (from the raster backward's slots) and the
// radii copy; in a capture the camera and Adam scalars come from the device StepParams
int eager_step_tail(msplat_hip_ctx *ctx, const StepRecord &r) {
    BackwardArgs b;
    CHECKED(backward_args(ctx, &r.g, &r.cam, b));
    for (int i = 0; i < 6; i++) {
        b.m1[i] = r.adam.exp_avg[i];
        b.m2[i] = r.adam.exp_avg_sq[i];
        b.step[i] = r.adam.step_size[i];
        b.bc2[i] = r.adam.bc2_sqrt[i];
    }
    b.beta1 = r.adam.beta1;
    b.beta2 = r.adam.beta2;
    b.eps = r.adam.eps;
    b.vis = r.stats.vis_counts;
    b.gn = r.stats.xys_grad_norm;
    b.ms = r.stats.max_2d_size;
    b.inv_max_dim = r.inv_max;
    b.sp = ctx->capturing ? (const StepParams *)ctx->step_params.p : nullptr;
    b.overflow = ctx->capturing ? ctx->d_overflow() : nullptr;
    ctx->begin(5);
    launch_backward(b, BWD_ADAM, ctx->stream);
    ctx->end();
    HIPCHECK(hipGetLastError());
    if (r.radii_out && r.g.num_points > 0)
        HIPCHECK(hipMemcpyAsync(r.radii_out, ctx->radii.p, 4 * (size_t)r.g.num_points, hipMemcpyDeviceToDevice,
                                ctx->stream));
    return MSPLAT_OK;
}

// One train step, launch by launch (the counts are read back to size the key buffers)
int eager_step(msplat_hip_ctx *ctx, const StepRecord &r, float *loss_out) {
    CHECKED(do_project(ctx, &r.g, &r.cam));
    CHECKED(do_bin(ctx));
    CHECKED(do_raster_forward(ctx, r.bg));
    CHECKED(do_loss(ctx, r.gt, r.ssim_weight, r.inv_n));
    CHECKED(do_raster_backward(ctx));
    CHECKED(eager_step_tail(ctx, r));
    if (ctx->step_graph_on) ctx->step_m_hint = ctx->M;
    if (loss_out) {
        HIPCHECK(hipMemcpyAsync(loss_out, ctx->d_loss(), 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
    }
    return MSPLAT_OK;
}

// ── the captured one-view train step ──────────────────────────────────────────
// A graph holds one step for one shape: gaussian count and buffers, view size, SH
// degree, background, loss weights, Adam constants, tuning. The per-step values
// (camera pose, SH bands in use, Adam step sizes and bias corrections, target) travel in
// a device StepParams block uploaded before each replay. The intersection buffers have a
// fixed capacity: a replay over capacity sets a sticky overflow word, sees an empty tile
// list and changes no parameter, moment or statistic, and every later replay does the
// same until the host clears the word. The host keeps the inputs of the replays it has
// not yet seen complete (STEP_RING of them): when it finds the word set, it waits for the
// stream, runs the replays that were not applied again, launch by launch and in order,
// raises the capacity and recaptures. The result is the same sequence of steps as the
// launch-by-launch path, with no host wait inside a step.

std::vector<uint64_t> shape_key(msplat_hip_ctx *ctx, const StepRecord &r) {
    auto bits = [](float f) {
        uint32_t u;
        memcpy(&u, &f, 4);
        return (uint64_t)u;
    };
    std::vector<uint64_t> k = {(uint64_t)r.g.num_points, r.cam.width, r.cam.height, r.cam.degree,
                               bits(r.bg[0]), bits(r.bg[1]), bits(r.bg[2]), bits(r.ssim_weight), bits(r.inv_n),
                               bits(r.inv_max), bits(r.adam.beta1), bits(r.adam.beta2), bits(r.adam.eps),
                               (uint64_t)ctx->flags, (uint64_t)ctx->tuning.chunks, (uint64_t)ctx->tuning.bwd_ppt,
                               (uint64_t)ctx->tuning.fwd_batch, (uint64_t)ctx->tuning.bwd_batch,
                               (uint64_t)ctx->tuning.cull, (uint64_t)ctx->tuning.pack_keys,
                               (uint64_t)(uintptr_t)r.radii_out};
    const void *ptrs[] = {r.g.means, r.g.scales, r.g.quats, r.g.features_dc, r.g.features_rest, r.g.opacities,
                          r.stats.vis_counts, r.stats.xys_grad_norm, r.stats.max_2d_size};
    for (const void *p : ptrs) k.push_back((uint64_t)(uintptr_t)p);
    for (int i = 0; i < 6; i++) {
        k.push_back((uint64_t)(uintptr_t)r.adam.exp_avg[i]);
        k.push_back((uint64_t)(uintptr_t)r.adam.exp_avg_sq[i]);
    }
    return k;
}

void drop_step_graph(msplat_hip_ctx *ctx) {
    if (ctx->step_exec) (void)hipGraphExecDestroy(ctx->step_exec);
    if (ctx->step_graph) (void)hipGraphDestroy(ctx->step_graph);
    ctx->step_exec = nullptr;
    ctx->step_graph = nullptr;
}

// Depth chunks of a frame with m intersections on the context's current view
int frame_chunks_for(msplat_hip_ctx *ctx, int64_t m) {
    const int tc = ctx->tuning.chunks;
    return m <= 0 ? 1 : tc == 0 ? auto_chunks(ctx->ntiles, m) : tc;
}

int capture_step(msplat_hip_ctx *ctx, const StepRecord &r) {
    drop_step_graph(ctx);
    // the capacity's buffers; every other one was sized by this shape's last eager step
    const size_t Mc = (size_t)ctx->step_cap, N = (size_t)(r.g.num_points > 0 ? r.g.num_points : 1);
    CHECKED(ctx->grow(ctx->keys0, 4 * Mc));
    CHECKED(ctx->grow(ctx->keys1, 4 * Mc));
    CHECKED(ctx->grow(ctx->vals0, 4 * Mc));
    CHECKED(ctx->grow(ctx->vals1, 4 * Mc));
    CHECKED(ctx->grow(ctx->hist, 4 * radix_hist_words((uint32_t)std::max(Mc, N))));
    CHECKED(ctx->grow(ctx->slot_f, (Mc + 3) & ~(size_t)3));
    CHECKED(ctx->grow(ctx->slot_a, 16 * Mc));
    CHECKED(ctx->grow(ctx->slot_b, 16 * Mc));
    CHECKED(ctx->grow(ctx->slot_c, 4 * Mc));
    ctx->ntiles = r.cam.tiles_x * r.cam.tiles_y;
    const int K = frame_chunks_for(ctx, ctx->step_m_hint);
    if (K > 1) {
        const size_t planes = (size_t)K * r.cam.width * r.cam.height;
        CHECKED(ctx->grow(ctx->chunk_prod, 4 * planes));
        CHECKED(ctx->grow(ctx->chunk_state, 16 * planes));
        CHECKED(ctx->grow(ctx->chunk_last, 4 * planes));
    }
    CHECKED(ctx->grow(ctx->step_params, sizeof(StepParams)));
    if (!ctx->h_params)
        HIPCHECK(hipHostMalloc((void **)&ctx->h_params, sizeof(StepParams) * STEP_RING, hipHostMallocDefault));
    if (!ctx->h_flag) {
        HIPCHECK(hipHostMalloc((void **)&ctx->h_flag, 64, hipHostMallocDefault));
        *ctx->h_flag = 0;
    }
    for (hipEvent_t &e : ctx->ring_ev)
        if (!e) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ctx->M = ctx->step_m_hint;  // the chunk choice inside the capture
    HIPCHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    const StepParams *sp = (const StepParams *)ctx->step_params.p;
    int rc = do_project(ctx, &r.g, &r.cam, sp);
    if (rc == MSPLAT_OK) rc = do_bin(ctx, (uint32_t)ctx->step_cap);
    if (rc == MSPLAT_OK) rc = do_raster_forward(ctx, r.bg);
    if (rc == MSPLAT_OK) rc = do_loss(ctx, r.gt, r.ssim_weight, r.inv_n, sp);
    if (rc == MSPLAT_OK) rc = do_raster_backward(ctx);
    if (rc == MSPLAT_OK) rc = eager_step_tail(ctx, r);
    const hipError_t launch_err = hipGetLastError();
    if (rc == MSPLAT_OK && launch_err == hipSuccess)
        (void)hipMemcpyAsync(ctx->h_flag, ctx->d_overflow(), 4, hipMemcpyDeviceToHost, ctx->stream);
    ctx->capturing = false;
    hipGraph_t graph = nullptr;
    const hipError_t end_err = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != MSPLAT_OK || launch_err != hipSuccess || end_err != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        ctx->frame_stage = msplat_hip_ctx::FRAME_NONE;
        if (rc != MSPLAT_OK) return rc;
        return set_error(MSPLAT_ERR_HIP, "train-step capture failed: %s",
                         hipGetErrorString(launch_err != hipSuccess ? launch_err : end_err));
    }
    ctx->step_graph = graph;
    HIPCHECK(hipGraphInstantiate(&ctx->step_exec, graph, nullptr, nullptr, 0));
    ctx->step_gen = ctx->generation;
    ctx->step_captures++;
    ctx->step_frame = StepFrame{ctx->n,           ctx->width,      ctx->height,        ctx->tiles_x,
                                ctx->tiles_y,     ctx->ntiles,     ctx->keys_sorted,   ctx->gid_sorted,
                                ctx->sort_gid_bits, ctx->vis_sorted, ctx->frame_chunks, {ctx->bg[0], ctx->bg[1], ctx->bg[2]}};
    return MSPLAT_OK;
}

// The frame state a replay leaves: the captured step's
void restore_step_frame(msplat_hip_ctx *ctx) {
    const StepFrame &f = ctx->step_frame;
    ctx->n = f.n;
    ctx->width = f.width;
    ctx->height = f.height;
    ctx->tiles_x = f.tiles_x;
    ctx->tiles_y = f.tiles_y;
    ctx->ntiles = f.ntiles;
    ctx->keys_sorted = f.keys_sorted;
    ctx->gid_sorted = f.gid_sorted;
    ctx->sort_gid_bits = f.sort_gid_bits;
    ctx->vis_sorted = f.vis_sorted;
    ctx->frame_chunks = f.chunks;
    for (int i = 0; i < 3; i++) ctx->bg[i] = f.bg[i];
    ctx->frame_stage = msplat_hip_ctx::FRAME_RASTERED;
    ctx->slot_f_clean = false;
    ctx->frame_aux = false;
}

// Waits for the replays in flight; runs again, launch by launch, those an overflow left
// unapplied, raises the capacity for them and drops the graph. Leaves the frame state of
// the last step.
int settle_steps(msplat_hip_ctx *ctx) {
    if (ctx->pending == 0) return MSPLAT_OK;
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    const int pending = ctx->pending;
    ctx->pending = 0;
    HIPCHECK(hipMemcpy(ctx->h_total, ctx->scalars.p, 128, hipMemcpyDeviceToHost));
    const uint32_t *w = (const uint32_t *)ctx->h_total;
    const uint32_t over = w[16], applied = w[17];
    ctx->M = (int64_t)ctx->h_total[0];
    ctx->V = (int64_t)ctx->h_total[5];
    if (!over) return MSPLAT_OK;
    const uint32_t unapplied = (uint32_t)ctx->issued - applied;
    zero_u32<<<1, 64, 0, ctx->stream>>>(ctx->d_overflow(), 2);
    HIPCHECK(hipGetLastError());
    *ctx->h_flag = 0;
    ctx->issued = 0;
    if (unapplied == 0 || unapplied > (uint32_t)pending)
        return set_error(MSPLAT_ERR_HIP, "internal: %u unapplied replays of %d in flight", unapplied, pending);
    drop_step_graph(ctx);
    int64_t mmax = 0;
    for (uint32_t i = unapplied; i > 0; i--) {
        const StepRecord &r = ctx->ring[(ctx->next_slot - i) % STEP_RING];
        CHECKED(eager_step(ctx, r, nullptr));
        mmax = std::max(mmax, ctx->M);
    }
    ctx->step_reruns += unapplied;
    ctx->step_cap = std::max<uint64_t>(ctx->step_cap, (uint64_t)mmax + (uint64_t)mmax / 4 + 4096);
    return MSPLAT_OK;
}

int graph_step(msplat_hip_ctx *ctx, const StepRecord &r, float *loss_out) {
    if (!ctx->stream) return set_error(MSPLAT_ERR_ARG, "graph capture needs a non-null stream");
    if (ctx->pending && *ctx->h_flag) CHECKED(settle_steps(ctx));  // a replay went over capacity
    const std::vector<uint64_t> shape = shape_key(ctx, r);
    if (!ctx->step_cap || shape != ctx->step_key) {
        // a new shape: this step runs launch by launch, sizing the buffers and the capacity
        CHECKED(settle_steps(ctx));
        drop_step_graph(ctx);
        CHECKED(eager_step(ctx, r, loss_out));
        const uint64_t m = (uint64_t)std::max<int64_t>(ctx->M, 0);
        ctx->step_cap = ctx->step_cap_request ? std::max<uint64_t>(ctx->step_cap_request, ctx->step_cap)
                                              : m + m / 4 + 4096;
        ctx->step_key = shape;
        return MSPLAT_OK;
    }
    const uint64_t slot = ctx->next_slot % STEP_RING;
    if (ctx->pending == STEP_RING) {  // the slot's replay: wait for it and check it applied
        HIPCHECK(hipEventSynchronize(ctx->ring_ev[slot]));
        if (*ctx->h_flag) CHECKED(settle_steps(ctx));  // it (or one after it) was over capacity
        else ctx->pending--;
    }
    if (!ctx->step_exec || ctx->step_gen != ctx->generation) {
        CHECKED(settle_steps(ctx));
        CHECKED(capture_step(ctx, r));
    }
    StepParams &p = ctx->h_params[slot];
    p.cam = r.cam;
    for (int i = 0; i < 6; i++) {
        p.step[i] = r.adam.step_size[i];
        p.bc2[i] = r.adam.bc2_sqrt[i];
    }
    p.gt = r.gt;
    ctx->ring[slot] = r;
    HIPCHECK(hipMemcpyAsync(ctx->step_params.p, &p, sizeof(StepParams), hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipGraphLaunch(ctx->step_exec, ctx->stream));
    HIPCHECK(hipEventRecord(ctx->ring_ev[slot], ctx->stream));
    ctx->next_slot++;
    ctx->pending++;
    ctx->issued++;
    ctx->step_replays++;
    restore_step_frame(ctx);
    if (loss_out) {
        CHECKED(settle_steps(ctx));  // a re-run (if any) ends with this step, whose loss is then the last
        HIPCHECK(hipMemcpy(loss_out, ctx->d_loss(), 4, hipMemcpyDeviceToHost));
    }
    return MSPLAT_OK;
}

// ── prior-guided frames (PriorStep) ───────────────────────────────────────────
int check_prior(const msplat_hip_prior *p, bool train) {
    if (!p) return set_error(MSPLAT_ERR_ARG, "prior is NULL");
    if (p->aux && p->sky && p->sky_tex && (p->sky_w <= 0 || p->sky_h <= 0))
        return set_error(MSPLAT_ERR_ARG, "sky texture size must be positive");
    if (train && p->aux && p->sky && p->sky_tex && p->sky_grad && (!p->sky_exp_avg || !p->sky_exp_avg_sq))
        return set_error(MSPLAT_ERR_ARG, "sky Adam moments are NULL");
    if (p->aux && (p->prior_depth || p->prior_aux) &&
        (!p->prior_depth || !p->prior_aux || p->prior_w <= 0 || p->prior_h <= 0))
        return set_error(MSPLAT_ERR_ARG, "a prior needs both its depth and aux images and a positive size");
    if (train && p->exposure && p->expo_params) {
        if (!p->expo_exp_avg || !p->expo_exp_avg_sq) return set_error(MSPLAT_ERR_ARG, "exposure moments are NULL");
        if (p->num_cams <= 0 || std::max(0, p->cam_index) >= p->num_cams)
            return set_error(MSPLAT_ERR_ARG, "exposure cam_index %d outside [0, %d)", p->cam_index, p->num_cams);
    }
    return MSPLAT_OK;
}

SkyArgs sky_args(const msplat_hip_prior *p, const msplat_hip_camera *cam, bool on) {
    SkyArgs k;
    memset(&k, 0, sizeof k);
    for (int i = 0; i < 3; i++) {
        k.up[i] = p->sky_frame[i];
        k.e1[i] = p->sky_frame[3 + i];
        k.e2[i] = p->sky_frame[6 + i];
    }
    k.fx = cam->fx;
    k.fy = cam->fy;
    k.cx = cam->cx;
    k.cy = cam->cy;
    for (int i = 0; i < 12; i++) k.vm[i] = cam->viewmat[i];
    k.tex = on ? p->sky_tex : nullptr;
    k.w = on ? (uint32_t)p->sky_w : 1u;
    k.h = on ? (uint32_t)p->sky_h : 1u;
    return k;
}

// An aux frame after its projection: the per-pixel background (the learned sky where
// it is on) and the aux buffers
int prepare_aux(msplat_hip_ctx *ctx, const msplat_hip_prior *p, const msplat_hip_camera *cam, const float bg[3]) {
    ctx->frame_aux = true;
    const size_t P = (size_t)cam->width * cam->height;
    CHECKED(ctx->grow(ctx->bg_img, 12 * P));
    CHECKED(ctx->grow(ctx->out_depth, 4 * P));
    CHECKED(ctx->grow(ctx->v_aux, 8 * P));
    launch_sky_background(sky_args(p, cam, p->sky && p->sky_tex), cam->width, cam->height, bg, (float *)ctx->bg_img.p,
                          ctx->stream);
    HIPCHECK(hipGetLastError());
    return MSPLAT_OK;
}

// msplat_train_step with a PriorStep (cpu:1219-1327), launch by launch
int prior_step(msplat_hip_ctx *ctx, const StepRecord &r, const msplat_hip_prior *p, float *loss_out) {
    const bool aux = p->aux != 0;
    const bool sky = aux && p->sky && p->sky_tex && p->sky_grad;
    const bool exposure = p->exposure && p->expo_params;
    const msplat_hip_camera *cam = &r.cam;
    const uint32_t W = cam->width, H = cam->height;
    const size_t P = (size_t)W * H;
    CHECKED(do_project(ctx, &r.g, cam));
    if (aux) CHECKED(prepare_aux(ctx, p, cam, r.bg));
    CHECKED(do_bin(ctx));
    CHECKED(do_raster_forward(ctx, r.bg));
    hipStream_t s = ctx->stream;
    // the photometric loss, on the exposure-adjusted render with exposure
    float *loss_grad = (float *)ctx->v_image.p;
    if (exposure) {
        CHECKED(ctx->grow(ctx->adj, 12 * P));
        CHECKED(ctx->grow(ctx->v_adj, 12 * P));
        launch_exposure_apply((const float *)ctx->image.p, p->expo_params + 12 * (size_t)std::max(0, p->cam_index), P,
                              (float *)ctx->adj.p, s);
        loss_grad = (float *)ctx->v_adj.p;
        CHECKED(do_loss(ctx, r.gt, r.ssim_weight, r.inv_n, nullptr, (const float *)ctx->adj.p, loss_grad));
    } else {
        CHECKED(do_loss(ctx, r.gt, r.ssim_weight, r.inv_n));
    }
    if (aux) {  // prior losses: dL/dD, dL/dA; the photometric gradient masked
        CHECKED(ctx->grow(ctx->prior_rows, 24 * (size_t)H));
        PriorLossArgs a;
        memset(&a, 0, sizeof a);
        a.width = W;
        a.height = H;
        a.final_T = (const float *)ctx->final_T.p;
        a.out_depth = (const float *)ctx->out_depth.p;
        a.v_photo = loss_grad;
        a.v_aux = (float *)ctx->v_aux.p;
        a.rows = (double *)ctx->prior_rows.p;
        const bool has_prior = p->prior_depth && p->prior_aux && p->prior_w > 0 && p->prior_h > 0;
        a.prior_depth = has_prior ? p->prior_depth : nullptr;
        a.prior_aux = has_prior ? p->prior_aux : nullptr;
        a.pw = has_prior ? (uint32_t)p->prior_w : 1u;
        a.ph = has_prior ? (uint32_t)p->prior_h : 1u;
        a.has_depth = has_prior && p->has_depth;
        a.has_sky = has_prior && p->has_sky_mask;
        a.depth_weight = p->depth_weight;
        a.sky_weight = p->sky_weight;
        a.fill_weight = p->fill_weight;
        a.huber = std::max(p->huber_delta, 1e-6f);
        a.min_alpha = p->min_alpha;
        a.mask_photometric = p->mask_photometric;
        launch_prior_loss(a, p->loss_terms, s);
    }
    if (exposure) {  // v_image = M^T v_adj, and the exposure row's Adam step
        CHECKED(ctx->grow(ctx->expo_part, 96 * std::max<size_t>(exposure_partials(P), 1)));
        launch_exposure_backward_adam((const float *)ctx->image.p, (const float *)ctx->v_adj.p, P,
                                      (float *)ctx->v_image.p, (double *)ctx->expo_part.p, p->expo_params,
                                      p->expo_exp_avg, p->expo_exp_avg_sq, p->expo_grad, std::max(0, p->cam_index),
                                      p->expo_step_size, p->expo_bc2_sqrt, p->expo_reg, r.adam.beta1, r.adam.beta2,
                                      r.adam.eps, s);
    }
    if (sky) {
        CHECKED(ctx->grow(ctx->sky_acc, 24 * (size_t)p->sky_w * p->sky_h));
        launch_sky_backward_adam(sky_args(p, cam, true), W, H, (const float *)ctx->final_T.p,
                                 (const float *)ctx->v_image.p, (unsigned long long *)ctx->sky_acc.p, p->sky_grad,
                                 p->sky_tex, p->sky_exp_avg, p->sky_exp_avg_sq, p->sky_step_size, p->sky_bc2_sqrt,
                                 r.adam.beta1, r.adam.beta2, r.adam.eps, s);
    }
    HIPCHECK(hipGetLastError());
    CHECKED(do_raster_backward(ctx));
    CHECKED(eager_step_tail(ctx, r));
    if (p->log_max_scale_ratio > 0.f) launch_scale_cap(r.g.scales, r.g.num_points, p->log_max_scale_ratio, s);
    HIPCHECK(hipGetLastError());
    if (loss_out) {
        HIPCHECK(hipMemcpyAsync(loss_out, ctx->d_loss(), 4, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
    }
    return MSPLAT_OK;
}

}  // namespace

int msplat::settle_replays(msplat_hip_ctx *ctx) { return settle_steps(ctx); }

// ═══════════════════════════════ public API ════════════════════════════════

extern "C" {

const char *msplat_hip_last_error(void) { return g_last_error.c_str(); }

int msplat_hip_ctx_create(int device, void *stream, msplat_hip_ctx **out) {
    if (!out) return set_error(MSPLAT_ERR_ARG, "out is NULL");
    HIPCHECK(hipSetDevice(device));
    auto *ctx = new msplat_hip_ctx();
    ctx->device = device;
    if (stream) {
        ctx->stream = (hipStream_t)stream;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete ctx;
            return set_error(MSPLAT_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
        }
        ctx->own_stream = true;
    }
    *out = ctx;
    return MSPLAT_OK;
}

int msplat_hip_ctx_destroy(msplat_hip_ctx *ctx) {
    if (ctx) (void)settle_steps(ctx);
    delete ctx;
    return MSPLAT_OK;
}

int msplat_hip_ctx_set_stream(msplat_hip_ctx *ctx, void *stream) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    if (ctx->own_stream) HIPCHECK(hipStreamDestroy(ctx->stream));
    ctx->own_stream = false;
    ctx->stream = (hipStream_t)stream;
    return MSPLAT_OK;
}

void *msplat_hip_ctx_stream(msplat_hip_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int msplat_hip_sync(msplat_hip_ctx *ctx) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    return MSPLAT_OK;
}

int msplat_hip_commit(msplat_hip_ctx *ctx) { return ctx ? MSPLAT_OK : set_error(MSPLAT_ERR_ARG, "ctx is NULL"); }

int msplat_hip_malloc(size_t bytes, int zero, void **out) {
    if (!out) return set_error(MSPLAT_ERR_ARG, "out is NULL");
    HIPCHECK(hipMalloc(out, bytes ? bytes : 4));
    if (zero) {
        HIPCHECK(hipMemset(*out, 0, bytes ? bytes : 4));
        HIPCHECK(hipDeviceSynchronize());
    }
    return MSPLAT_OK;
}

int msplat_hip_device_memory(int device, size_t *free_bytes, size_t *total_bytes) {
    if (!free_bytes || !total_bytes) return set_error(MSPLAT_ERR_ARG, "NULL output");
    int prev = 0;
    HIPCHECK(hipGetDevice(&prev));
    HIPCHECK(hipSetDevice(device));
    const hipError_t e = hipMemGetInfo(free_bytes, total_bytes);
    HIPCHECK(hipSetDevice(prev));
    if (e != hipSuccess) return set_error(MSPLAT_ERR_HIP, "hipMemGetInfo failed: %s", hipGetErrorString(e));
    return MSPLAT_OK;
}

int msplat_hip_free(void *ptr) {
    if (ptr) HIPCHECK(hipFree(ptr));
    return MSPLAT_OK;
}

int msplat_hip_memcpy(msplat_hip_ctx *ctx, void *dst, const void *src, size_t bytes, int kind) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    if (bytes == 0) return MSPLAT_OK;
    if (kind == 0) {
        HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));  // the host buffer may be reused on return
    } else if (kind == 1) {
        HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
    } else if (kind == 2) {
        HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
        return set_error(MSPLAT_ERR_ARG, "unknown copy kind %d", kind);
    }
    return MSPLAT_OK;
}

int msplat_hip_memset(msplat_hip_ctx *ctx, void *dst, int value, size_t bytes) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    if (bytes) HIPCHECK(hipMemsetAsync(dst, value, bytes, ctx->stream));
    return MSPLAT_OK;
}

int msplat_hip_stage_project(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_camera *cam) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    CHECKED(check_gaussians(g));
    CHECKED(check_camera(cam));
    return do_project(ctx, g, cam);
}

int msplat_hip_stage_bin(msplat_hip_ctx *ctx, const msplat_hip_camera *cam) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    CHECKED(check_camera(cam));
    if (ctx->frame_stage < msplat_hip_ctx::FRAME_PROJECTED)
        return set_error(MSPLAT_ERR_ARG, "binning needs a projected frame on this context");
    if (cam->tiles_x != ctx->tiles_x || cam->tiles_y != ctx->tiles_y)
        return set_error(MSPLAT_ERR_ARG, "camera differs from the projected frame");
    // rectangles, counts and per-block sums are re-derived from the frame's projection
    // outputs, so that a caller may overwrite them before binning
    launch_rect_count(ctx->n, (const float4 *)ctx->splat_a.p, (const float2 *)ctx->aabb.p,
                      (const int32_t *)ctx->radii.p, ctx->tiles_x, ctx->tiles_y, (uint2 *)ctx->rect.p,
                      (uint32_t *)ctx->count.p, (uint32_t *)ctx->block_sums.p, (uint32_t *)ctx->block_vis.p,
                      ctx->stream);
    HIPCHECK(hipGetLastError());
    return do_bin(ctx);
}

int msplat_hip_stage_raster_forward(msplat_hip_ctx *ctx, const msplat_hip_camera *cam, const float background[3]) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    CHECKED(check_camera(cam));
    if (ctx->frame_stage >= msplat_hip_ctx::FRAME_PROJECTED &&
        (cam->width != ctx->width || cam->height != ctx->height))
        return set_error(MSPLAT_ERR_ARG, "camera differs from the binned frame");
    return do_raster_forward(ctx, background);
}

int msplat_hip_stage_loss(msplat_hip_ctx *ctx, const msplat_hip_camera *cam, const float *gt, float ssim_weight,
                          float loss_inv_n, float *loss_out) {
    if (!ctx || !gt) return set_error(MSPLAT_ERR_ARG, "ctx or gt is NULL");
    CHECKED(settle_steps(ctx));
    CHECKED(check_camera(cam));
    CHECKED(do_loss(ctx, gt, ssim_weight, loss_inv_n));
    if (loss_out) {
        HIPCHECK(hipMemcpyAsync(loss_out, ctx->d_loss(), 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHECK(hipStreamSynchronize(ctx->stream));
    }
    return MSPLAT_OK;
}

int msplat_hip_render(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_camera *cam,
                      const float background[3], float **out_img) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    CHECKED(check_gaussians(g));
    CHECKED(check_camera(cam));
    CHECKED(do_project(ctx, g, cam));
    CHECKED(do_bin(ctx));
    CHECKED(do_raster_forward(ctx, background));
    if (out_img) *out_img = (float *)ctx->image.p;
    return MSPLAT_OK;
}

int msplat_hip_train_step(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_camera *cam,
                          const float background[3], const float *gt, float ssim_weight, float loss_inv_n,
                          const msplat_hip_adam *adam, const msplat_hip_grad_stats *stats, float inv_max_dim,
                          int32_t *radii_out, float *loss_out) {
    if (!ctx || !gt || !adam || !stats || !background)
        return set_error(MSPLAT_ERR_ARG, "ctx, background, gt, adam or stats is NULL");
    CHECKED(check_gaussians(g));
    CHECKED(check_camera(cam));
    for (int i = 0; i < 6; i++)
        if (g->num_points > 0 && (!adam->exp_avg[i] || !adam->exp_avg_sq[i]))
            return set_error(MSPLAT_ERR_ARG, "Adam moment %d is NULL", i);
    if (g->num_points > 0 && cam->degree > 0 && !g->features_rest)
        return set_error(MSPLAT_ERR_ARG, "features_rest is NULL");
    StepRecord r;
    r.g = *g;
    r.cam = *cam;
    for (int i = 0; i < 3; i++) r.bg[i] = background[i];
    r.gt = gt;
    r.ssim_weight = ssim_weight;
    r.inv_n = loss_inv_n;
    r.inv_max = inv_max_dim;
    r.adam = *adam;
    r.stats = *stats;
    r.radii_out = radii_out;
    if (ctx->step_graph_on) return graph_step(ctx, r, loss_out);
    CHECKED(settle_steps(ctx));
    return eager_step(ctx, r, loss_out);
}

int msplat_hip_train_step_prior(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_camera *cam,
                                const float background[3], const float *gt, float ssim_weight, float loss_inv_n,
                                const msplat_hip_adam *adam, const msplat_hip_grad_stats *stats, float inv_max_dim,
                                int32_t *radii_out, float *loss_out, const msplat_hip_prior *prior) {
    if (!ctx || !gt || !adam || !stats || !background)
        return set_error(MSPLAT_ERR_ARG, "ctx, background, gt, adam or stats is NULL");
    CHECKED(settle_steps(ctx));
    CHECKED(check_gaussians(g));
    CHECKED(check_camera(cam));
    CHECKED(check_prior(prior, true));
    for (int i = 0; i < 6; i++)
        if (g->num_points > 0 && (!adam->exp_avg[i] || !adam->exp_avg_sq[i]))
            return set_error(MSPLAT_ERR_ARG, "Adam moment %d is NULL", i);
    if (g->num_points > 0 && cam->degree > 0 && !g->features_rest)
        return set_error(MSPLAT_ERR_ARG, "features_rest is NULL");
    StepRecord r;
    r.g = *g;
    r.cam = *cam;
    for (int i = 0; i < 3; i++) r.bg[i] = background[i];
    r.gt = gt;
    r.ssim_weight = ssim_weight;
    r.inv_n = loss_inv_n;
    r.inv_max = inv_max_dim;
    r.adam = *adam;
    r.stats = *stats;
    r.radii_out = radii_out;
    return prior_step(ctx, r, prior, loss_out);
}

int msplat_hip_render_prior(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_camera *cam,
                            const float background[3], const msplat_hip_prior *prior, float **out_img) {
    if (!ctx || !background) return set_error(MSPLAT_ERR_ARG, "ctx or background is NULL");
    CHECKED(settle_steps(ctx));
    CHECKED(check_gaussians(g));
    CHECKED(check_camera(cam));
    CHECKED(check_prior(prior, false));
    CHECKED(do_project(ctx, g, cam));
    if (prior->aux) CHECKED(prepare_aux(ctx, prior, cam, background));
    CHECKED(do_bin(ctx));
    CHECKED(do_raster_forward(ctx, background));
    if (out_img) *out_img = (float *)ctx->image.p;
    return MSPLAT_OK;
}

int msplat_hip_render_aux_outputs(msplat_hip_ctx *ctx, float **depth, float **final_T) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    if (ctx->frame_stage < msplat_hip_ctx::FRAME_RASTERED || !ctx->frame_aux)
        return set_error(MSPLAT_ERR_ARG, "no prior-guided (aux) frame on this context");
    if (depth) *depth = (float *)ctx->out_depth.p;
    if (final_T) *final_T = (float *)ctx->final_T.p;
    return MSPLAT_OK;
}

int msplat_hip_finalize_depth(msplat_hip_ctx *ctx, const float *depth_num, const float *final_T, uint32_t n,
                              float inv_scale, float *depth_out, float *alpha_out) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    if (n && (!depth_num || !final_T || !depth_out || !alpha_out)) return set_error(MSPLAT_ERR_ARG, "NULL array");
    launch_finalize_depth(depth_num, final_T, n, inv_scale, depth_out, alpha_out, ctx->stream);
    HIPCHECK(hipGetLastError());
    return MSPLAT_OK;
}

int msplat_hip_set_step_graph(msplat_hip_ctx *ctx, int enable, uint64_t initial_capacity) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    if (enable && !ctx->stream) return set_error(MSPLAT_ERR_ARG, "graph capture needs a non-null stream");
    if (initial_capacity > 0x7FFFFFFFull) return set_error(MSPLAT_ERR_RANGE, "capacity above 2^31 intersections");
    CHECKED(settle_steps(ctx));
    drop_step_graph(ctx);
    ctx->step_graph_on = enable != 0;
    ctx->step_cap_request = (uint32_t)initial_capacity;
    ctx->step_cap = 0;  // the next step runs launch by launch and sizes the capacity
    ctx->step_key.clear();
    return MSPLAT_OK;
}

int msplat_hip_step_graph_stats(msplat_hip_ctx *ctx, int64_t *captures, int64_t *replays, int64_t *reruns) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    if (captures) *captures = ctx->step_captures;
    if (replays) *replays = ctx->step_replays;
    if (reruns) *reruns = ctx->step_reruns;
    return MSPLAT_OK;
}

int msplat_hip_last_loss(msplat_hip_ctx *ctx, float *loss) {
    if (!ctx || !loss) return set_error(MSPLAT_ERR_ARG, "ctx or loss is NULL");
    CHECKED(settle_steps(ctx));
    if (!ctx->scalars.p) return set_error(MSPLAT_ERR_ARG, "no train step has run on this context");
    HIPCHECK(hipMemcpyAsync(loss, ctx->d_loss(), 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    return MSPLAT_OK;
}

int msplat_hip_render_backward(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_camera *cam,
                               const float background[3], const float *v_img, const msplat_hip_grads *out) {
    if (!ctx || !v_img || !out) return set_error(MSPLAT_ERR_ARG, "ctx, v_img or out is NULL");
    CHECKED(settle_steps(ctx));
    CHECKED(check_gaussians(g));
    CHECKED(check_camera(cam));
    if (ctx->frame_stage < msplat_hip_ctx::FRAME_RASTERED)
        return set_error(MSPLAT_ERR_ARG, "render_backward needs a render on this context first");
    if (g->num_points != ctx->n || cam->width != ctx->width || cam->height != ctx->height)
        return set_error(MSPLAT_ERR_ARG, "render_backward does not match the last render on this context");
    (void)background;
    const size_t P = (size_t)ctx->width * ctx->height;
    HIPCHECK(hipMemcpyAsync(ctx->v_image.p, v_img, 12 * P, hipMemcpyDeviceToDevice, ctx->stream));
    if (ctx->frame_aux)  // a prior-guided render: no prior-loss gradients, only the image's
        zero_u32<<<(unsigned)std::min<size_t>((2 * P + 255) / 256, 4096), 256, 0, ctx->stream>>>(
            (uint32_t *)ctx->v_aux.p, 2 * P);
    CHECKED(do_raster_backward(ctx));
    BackwardArgs b;
    CHECKED(backward_args(ctx, g, cam, b));
    const size_t N = (size_t)(g->num_points > 0 ? g->num_points : 1);
    const uint32_t nb = sh_bases(cam->degree);
    // NULL outputs go to a scratch area
    float *tmp = nullptr;
    size_t need = N * (3 + 3 + 4 + 3 + (nb - 1) * 3 + 1) * 4;
    if (!out->v_means || !out->v_scales || !out->v_quats || !out->v_features_dc || !out->v_features_rest ||
        !out->v_opacities) {
        CHECKED(ctx->grow(ctx->grad_scratch, need));
        tmp = (float *)ctx->grad_scratch.p;
    }
    float *cur = tmp;
    auto pick = [&](float *p, size_t count) {
        if (p) return p;
        float *r = cur;
        cur += count;
        return r;
    };
    b.g_means = pick(out->v_means, 3 * N);
    b.g_scales = pick(out->v_scales, 3 * N);
    b.g_quats = pick(out->v_quats, 4 * N);
    b.g_dc = pick(out->v_features_dc, 3 * N);
    b.g_rest = pick(out->v_features_rest, (nb - 1) * 3 * N);
    b.g_opac = pick(out->v_opacities, N);
    ctx->begin(5);
    launch_backward(b, BWD_GRADS, ctx->stream);
    ctx->end();
    HIPCHECK(hipGetLastError());
    return MSPLAT_OK;
}

int msplat_hip_frame_view(msplat_hip_ctx *ctx, msplat_hip_frame *f) {
    if (!ctx || !f) return set_error(MSPLAT_ERR_ARG, "ctx or out is NULL");
    CHECKED(settle_steps(ctx));
    memset(f, 0, sizeof *f);
    f->num_points = ctx->n;
    f->num_tiles = ctx->ntiles;
    f->num_isects = ctx->M;
    f->num_visible = ctx->V;
    f->depth_chunks = ctx->frame_chunks;
    f->splat_a = (float *)ctx->splat_a.p;
    f->splat_b = (float *)ctx->splat_b.p;
    f->splat_blue = (float *)ctx->splat_blue.p;
    f->depths = (float *)ctx->depths.p;
    f->aabb = (float *)ctx->aabb.p;
    f->radii = (int32_t *)ctx->radii.p;
    f->tile_count = (uint32_t *)ctx->count.p;
    f->offsets = (uint32_t *)ctx->offsets.p;
    f->isect_keys = nullptr;
    f->isect_gid = ctx->gid_sorted;
    if ((ctx->flags & MSPLAT_FRAME_KEYS64) && ctx->keys_sorted) {
        const uint32_t *tiles = ctx->keys_sorted;
        if (ctx->sort_gid_bits) {  // split the packed keys into the key buffer not holding them, and vals0
            uint32_t *spare = ctx->keys_sorted == (uint32_t *)ctx->keys0.p ? (uint32_t *)ctx->keys1.p
                                                                           : (uint32_t *)ctx->keys0.p;
            launch_unpack_keys(ctx->keys_sorted, (uint32_t)ctx->M, ctx->sort_gid_bits, spare,
                               (int32_t *)ctx->vals0.p, ctx->stream);
            tiles = spare;
            f->isect_gid = (int32_t *)ctx->vals0.p;
        }
        CHECKED(ctx->grow(ctx->keys64, 8 * (size_t)(ctx->M > 0 ? ctx->M : 1)));
        launch_keys64(tiles, f->isect_gid, (const float *)ctx->depths.p, (uint32_t)ctx->M,
                      (uint64_t *)ctx->keys64.p, ctx->stream);
        HIPCHECK(hipGetLastError());
        f->isect_keys = (uint64_t *)ctx->keys64.p;
    }
    f->tile_start = (int32_t *)ctx->tile_start.p;
    f->image = (float *)ctx->image.p;
    f->final_T = (float *)ctx->final_T.p;
    f->final_idx = (int32_t *)ctx->final_idx.p;
    f->v_image = (float *)ctx->v_image.p;
    f->loss_sum = ctx->scalars.p ? ctx->d_loss_sum() : nullptr;
    f->slots_written = ctx->scalars.p ? (uint64_t *)ctx->d_slots_written() : nullptr;
    f->isects_running = ctx->scalars.p ? ctx->d_isects_running() : nullptr;
    f->visible_running = ctx->scalars.p ? ctx->d_visible_running() : nullptr;
    f->v_xy = (float *)ctx->v_xy.p;
    f->v_conic = (float *)ctx->v_conic.p;
    f->v_rgb = (float *)ctx->v_rgb.p;
    f->v_opacity = (float *)ctx->v_opac.p;
    return MSPLAT_OK;
}

int msplat_hip_set_flags(msplat_hip_ctx *ctx, int flags) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    ctx->flags = flags;
    return MSPLAT_OK;
}

int msplat_hip_radix_sort(msplat_hip_ctx *ctx, uint64_t *keys, int32_t *vals, uint32_t n, int key_bits) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    if (key_bits < 0 || key_bits > 64) return set_error(MSPLAT_ERR_ARG, "key_bits must be in [0, 64]");
    if (n == 0 || key_bits == 0) return MSPLAT_OK;
    if (!keys || !vals) return set_error(MSPLAT_ERR_ARG, "keys or vals is NULL");
    CHECKED(ctx->grow(ctx->keys1, 8 * (size_t)n));
    CHECKED(ctx->grow(ctx->vals1, 4 * (size_t)n));
    CHECKED(ctx->grow(ctx->hist, 4 * radix_hist_words(n)));
    CHECKED(ctx->grow(ctx->digit_total, 4 * 256));
    CHECKED(ctx->grow(ctx->scalars, 64));
    uint32_t *dn = ctx->d_total32();
    fill_u32<<<1, 1, 0, ctx->stream>>>(dn, n);
    uint64_t *kin = keys, *kout = (uint64_t *)ctx->keys1.p;
    int32_t *vin = vals, *vout = (int32_t *)ctx->vals1.p;
    // only the low key_bits order the keys (higher bits ride along): the last pass is
    // narrowed to the bits left
    for (int shift = 0; shift < key_bits; shift += 8) {
        const int width = std::min(8, key_bits - shift);
        launch_radix_pass(kin, vin, kout, vout, dn, n, shift, (uint32_t *)ctx->hist.p, (uint32_t *)ctx->digit_total.p,
                          ctx->stream, (1u << width) - 1u);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (kin != keys) {
        HIPCHECK(hipMemcpyAsync(keys, kin, 8 * (size_t)n, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHECK(hipMemcpyAsync(vals, vin, 4 * (size_t)n, hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIPCHECK(hipGetLastError());
    // the frame's sorted intersections lived in keys1/vals1; they are no longer valid
    ctx->keys_sorted = nullptr;
    ctx->gid_sorted = nullptr;
    ctx->sort_gid_bits = 0;
    if (ctx->frame_stage > msplat_hip_ctx::FRAME_PROJECTED) ctx->frame_stage = msplat_hip_ctx::FRAME_PROJECTED;
    return MSPLAT_OK;
}

int msplat_hip_densify(msplat_hip_ctx *ctx, int32_t n, int32_t capacity, float grad_thresh, float size_thresh,
                       float screen_thresh, int check_screen, float cull_alpha_thresh, float cull_scale_thresh,
                       float cull_screen_size, int check_huge, const float *xys_grad_norm, const float *vis_counts,
                       const float *max_2d_size, float half_max_dim, float *const bufs[18], int32_t fr_stride,
                       uint32_t seed, int32_t *new_count) {
    if (!ctx || !bufs || !new_count) return set_error(MSPLAT_ERR_ARG, "ctx, bufs or new_count is NULL");
    CHECKED(settle_steps(ctx));
    if (n < 0 || fr_stride < 1) return set_error(MSPLAT_ERR_ARG, "bad gaussian count or rest stride");
    if ((int64_t)3 * n > capacity) return set_error(MSPLAT_ERR_CAPACITY, "msplat_densify: 3*N exceeds buf_capacity");
    for (int i = 0; i < 18; i++)
        if (n > 0 && !bufs[i]) return set_error(MSPLAT_ERR_ARG, "densify buffer %d is NULL", i);
    if (n == 0) {
        *new_count = 0;
        return MSPLAT_OK;
    }
    DensifyArgs a;
    memset(&a, 0, sizeof a);
    a.n = n;
    a.grad_thresh = grad_thresh;
    a.size_thresh = size_thresh;
    a.screen_thresh = screen_thresh;
    a.cull_alpha = cull_alpha_thresh;
    a.cull_scale = cull_scale_thresh;
    a.cull_screen = cull_screen_size;
    a.half_max_dim = half_max_dim;
    a.log_shrink = logf(1.6f);
    a.check_screen = check_screen;
    a.check_huge = check_huge;
    a.seed = seed;
    a.gn = xys_grad_norm;
    a.vc = vis_counts;
    a.ms = max_2d_size;
    for (int i = 0; i < 18; i++) a.bufs[i] = bufs[i];
    const int widths[6] = {3, 3, 4, 3, fr_stride, 1};
    for (int i = 0; i < 6; i++) a.width[i] = widths[i];
    const size_t N3 = 3 * (size_t)n, nb3 = (N3 + 255) / 256 + 1;
    // workspace: flags (3 x 3N bytes), block sums (3 x nb3), ranks (3N), scalars, scratch (3N x max width)
    const size_t maxw = (size_t)(fr_stride > 4 ? fr_stride : 4);
    const size_t bytes = 3 * N3 + 16 + 3 * 4 * nb3 + 4 * N3 + 64 + 4 * N3 * maxw + 256;
    CHECKED(ctx->grow(ctx->aux, bytes));
    char *p = (char *)ctx->aux.p;
    auto take = [&](size_t b) {
        char *r = p;
        p += (b + 15) & ~(size_t)15;
        return (void *)r;
    };
    a.split = (uint8_t *)take(N3);
    a.dup = (uint8_t *)take(N3);
    a.keep = (uint8_t *)take(N3);
    a.split_bs = (uint32_t *)take(4 * nb3);
    a.dup_bs = (uint32_t *)take(4 * nb3);
    a.keep_bs = (uint32_t *)take(4 * nb3);
    a.dst = (int32_t *)take(4 * N3);
    uint64_t *tot = (uint64_t *)take(64);
    float *scratch = (float *)take(4 * N3 * maxw);
    if (!ctx->h_total) HIPCHECK(hipHostMalloc((void **)&ctx->h_total, 64, hipHostMallocDefault));
    const int nb = (n + 255) / 256;
    launch_densify_classify(a, ctx->stream);
    launch_scan_blocks(a.split_bs, nb, tot, (uint32_t *)(tot + 3), ctx->stream);
    launch_scan_blocks(a.dup_bs, nb, tot + 1, (uint32_t *)(tot + 3) + 1, ctx->stream);
    HIPCHECK(hipMemcpyAsync(ctx->h_total, tot, 16, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    const int n_split = (int)ctx->h_total[0], n_dup = (int)ctx->h_total[1];
    a.n_split = n_split;
    a.n_new = n + 2 * n_split + n_dup;
    launch_densify_append(a, ctx->stream);
    launch_densify_cull(a, ctx->stream);
    const int nbn = (a.n_new + 255) / 256;
    launch_scan_blocks(a.keep_bs, nbn, tot + 2, (uint32_t *)(tot + 3) + 2, ctx->stream);
    launch_densify_rank(a, ctx->stream);
    HIPCHECK(hipMemcpyAsync(ctx->h_total + 2, tot + 2, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    const int kept = (int)ctx->h_total[2];
    for (int b = 0; b < 18; b++) {
        const int w = widths[b % 6];
        launch_compact_scatter(bufs[b], scratch, a.dst, a.n_new, w, ctx->stream);
        HIPCHECK(hipMemcpyAsync(bufs[b], scratch, 4 * (size_t)kept * w, hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    *new_count = kept;
    return MSPLAT_OK;
}

int msplat_hip_opacity_reset(msplat_hip_ctx *ctx, float *opacities, float *exp_avg, float *exp_avg_sq, int32_t n,
                             float reset_logit) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    if (n > 0 && (!opacities || !exp_avg || !exp_avg_sq)) return set_error(MSPLAT_ERR_ARG, "NULL buffer");
    launch_opacity_reset(opacities, exp_avg, exp_avg_sq, n, reset_logit, ctx->stream);
    HIPCHECK(hipGetLastError());
    return MSPLAT_OK;
}

int msplat_hip_knn3_mean_dist(msplat_hip_ctx *ctx, const float *points, uint32_t n, float *mean_dist) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    if (n == 0) return MSPLAT_OK;
    if (!points || !mean_dist) return set_error(MSPLAT_ERR_ARG, "points or mean_dist is NULL");
    CHECKED(ctx->grow(ctx->aux, knn_scratch_bytes(n) + 1024));
    KnnScratch w;
    knn_scratch_carve(ctx->aux.p, n, w);
    knn3_launch(points, n, mean_dist, w, ctx->stream);
    HIPCHECK(hipGetLastError());
    return MSPLAT_OK;
}

int msplat_hip_init_gaussians(msplat_hip_ctx *ctx, const float *means, const uint8_t *rgb, uint32_t n, uint32_t seed,
                              float opacity_logit, float *scales, float *quats, float *features_dc,
                              float *opacities) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    if (n == 0) return MSPLAT_OK;
    if (!means || !rgb || !scales || !quats || !features_dc || !opacities)
        return set_error(MSPLAT_ERR_ARG, "NULL buffer");
    // the mean distances live past the kNN workspace
    const size_t kb = (knn_scratch_bytes(n) + 1024 + 255) & ~(size_t)255;
    CHECKED(ctx->grow(ctx->aux, kb + 4 * (size_t)n));
    KnnScratch w;
    knn_scratch_carve(ctx->aux.p, n, w);
    float *md = (float *)((char *)ctx->aux.p + kb);
    knn3_launch(means, n, md, w, ctx->stream);
    launch_init_gaussians(md, rgb, n, seed, opacity_logit, scales, quats, features_dc, opacities, ctx->stream);
    HIPCHECK(hipGetLastError());
    return MSPLAT_OK;
}

int msplat_hip_image_metrics(msplat_hip_ctx *ctx, const float *rendered, const float *gt, int h, int w,
                             double out[3]) {
    if (!ctx || !out) return set_error(MSPLAT_ERR_ARG, "ctx or out is NULL");
    CHECKED(settle_steps(ctx));
    out[0] = out[1] = out[2] = 0.0;
    if (h <= 0 || w <= 0) return MSPLAT_OK;
    if (!rendered || !gt) return set_error(MSPLAT_ERR_ARG, "NULL image");
    const size_t P = (size_t)h * w;
    CHECKED(ctx->grow(ctx->aux, 4 * P * 15 + 4 * P * 3 + 8 * 3 * (size_t)h + 64));
    float *mom = (float *)ctx->aux.p, *ssim = mom + 15 * P;
    double *rows = (double *)(ssim + 3 * P), *res = rows + 3 * (size_t)h;  // 72 P bytes in: 8-aligned
    launch_metrics(rendered, gt, h, w, mom, ssim, rows, res, ctx->stream);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(out, res, 24, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    return MSPLAT_OK;
}

int msplat_hip_pack_rgba8(msplat_hip_ctx *ctx, const float *img, uint32_t n, uint8_t *out) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_steps(ctx));
    if (n == 0) return MSPLAT_OK;
    if (!img || !out) return set_error(MSPLAT_ERR_ARG, "img or out is NULL");
    launch_pack_rgba8(img, n, out, ctx->stream);
    HIPCHECK(hipGetLastError());
    return MSPLAT_OK;
}

int msplat_hip_set_tuning(msplat_hip_ctx *ctx, int key, int value) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    switch (key) {
    case MSPLAT_TUNE_BWD_PPT:
        if (value != 1 && value != 2 && value != 4)
            return set_error(MSPLAT_ERR_ARG, "backward pixels per thread must be 1, 2 or 4");
        ctx->tuning.bwd_ppt = value;
        return MSPLAT_OK;
    case MSPLAT_TUNE_CHUNKS:
        if (value < 0 || value > 64) return set_error(MSPLAT_ERR_ARG, "depth chunks must be 0 (per frame) or 1..64");
        ctx->tuning.chunks = value;
        return MSPLAT_OK;
    case MSPLAT_TUNE_FWD_BATCH:
        if (value != 64 && value != 128 && value != 256)
            return set_error(MSPLAT_ERR_ARG, "forward LDS batch depth must be 64, 128 or 256");
        ctx->tuning.fwd_batch = value;
        return MSPLAT_OK;
    case MSPLAT_TUNE_BWD_BATCH:
        if (value != 32 && value != 64 && value != 128)
            return set_error(MSPLAT_ERR_ARG, "backward LDS batch depth must be 32, 64 or 128");
        ctx->tuning.bwd_batch = value;
        return MSPLAT_OK;
    case MSPLAT_TUNE_CULL:
        if (value != 0 && value != 1) return set_error(MSPLAT_ERR_ARG, "cull must be 0 or 1");
        ctx->tuning.cull = value;
        return MSPLAT_OK;
    case MSPLAT_TUNE_PACK_KEYS:
        if (value != 0 && value != 1) return set_error(MSPLAT_ERR_ARG, "pack_keys must be 0 or 1");
        ctx->tuning.pack_keys = value;
        return MSPLAT_OK;
    default:
        return set_error(MSPLAT_ERR_ARG, "unknown tunable %d", key);
    }
}

int msplat_hip_get_tuning(msplat_hip_ctx *ctx, int key, int *value) {
    if (!ctx || !value) return set_error(MSPLAT_ERR_ARG, "ctx or value is NULL");
    switch (key) {
    case MSPLAT_TUNE_BWD_PPT: *value = ctx->tuning.bwd_ppt; return MSPLAT_OK;
    case MSPLAT_TUNE_FWD_BATCH: *value = ctx->tuning.fwd_batch; return MSPLAT_OK;
    case MSPLAT_TUNE_BWD_BATCH: *value = ctx->tuning.bwd_batch; return MSPLAT_OK;
    case MSPLAT_TUNE_CULL: *value = ctx->tuning.cull; return MSPLAT_OK;
    case MSPLAT_TUNE_PACK_KEYS: *value = ctx->tuning.pack_keys; return MSPLAT_OK;
    case MSPLAT_TUNE_CHUNKS: *value = ctx->tuning.chunks; return MSPLAT_OK;
    default: return set_error(MSPLAT_ERR_ARG, "unknown tunable %d", key);
    }
}

int msplat_hip_enable_stage_timing(msplat_hip_ctx *ctx, int enable) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->collect();
    ctx->timing = enable != 0;
    return MSPLAT_OK;
}

int msplat_hip_set_stage_mask(msplat_hip_ctx *ctx, uint32_t mask) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    if (mask >> MSPLAT_NUM_STAGES) return set_error(MSPLAT_ERR_ARG, "stage mask 0x%x names no stage", mask);
    ctx->stage_mask = mask;
    return MSPLAT_OK;
}

int msplat_hip_drain_stage_times(msplat_hip_ctx *ctx, double *ms, int64_t *calls, int max_stages, int *n_stages) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    ctx->collect();
    const int n = max_stages < MSPLAT_NUM_STAGES ? max_stages : MSPLAT_NUM_STAGES;
    for (int i = 0; i < n; i++) {
        if (ms) ms[i] = ctx->stage_ms[i];
        if (calls) calls[i] = ctx->stage_calls[i];
    }
    for (int i = 0; i < MSPLAT_NUM_STAGES; i++) {
        ctx->stage_ms[i] = 0.0;
        ctx->stage_calls[i] = 0;
    }
    if (n_stages) *n_stages = MSPLAT_NUM_STAGES;
    return MSPLAT_OK;
}

const char *msplat_hip_stage_name(int stage) {
    static const char *names[MSPLAT_NUM_STAGES] = {"proj_sh_fwd", "prefix_sort_pack", "rast_fwd",
                                                   "loss_fwd_bwd", "rast_bwd", "proj_sh_bwd_adam"};
    return stage >= 0 && stage < MSPLAT_NUM_STAGES ? names[stage] : "";
}

}  // extern "C"
