// Batched views (include/msplat_hip.h, "batched views"): one frame over several views
// of one scene, optionally replayed as a captured hipGraph.
//
// The frame's shape-independent inputs live in a BatchTable in device memory, uploaded
// before each launch; every kernel of the batch reads poses, resolutions, targets and
// Adam scalars from it, so a captured graph replays with new contents. Eager frames count
// the intersections first and size their buffers from the read-back totals. A captured
// step has fixed capacities instead: the count pass compares the total with them on the
// device, and over capacity every stage sees an empty frame and the Adam kernels do
// nothing, so the host grows the capacity and runs the step again (no partial update).
#include <algorithm>
#include <cstring>
#include <utility>

#include "context.h"

using namespace msplat;

namespace msplat {

struct BatchState {
    // per record (views x gaussians)
    Buf splat_a, splat_b, splat_blue, depths, rect, radii, count, offsets, block_sums, block_vis, big_list, big_sums;
    // visible records sorted by depth, count sums in depth order, scan partials
    Buf dk0, dk1, dv0, dv1, ord_sums, scan_part, loff;
    // per intersection
    Buf keys0, keys1, vals0, vals1, slot_a, slot_b, slot_c, slot_f, hist, digit_total;
    // per tile, pixel, row
    Buf tile_start, image, final_T, final_idx, v_image, row_loss, loss_scratch;
    // summed gradients (per gaussian), visibility in any view
    Buf grads, vis_any;
    Buf table, scalars;
    BatchTable *h_tab = nullptr;    // pinned staging of the table
    BatchScalars *h_sc = nullptr;   // pinned read-back of the scalars
    hipEvent_t tab_done = nullptr;  // the last table upload has been read
    bool tab_pending = false;
    uint32_t cap = 0;  // intersection capacity
    uint32_t tiles_cap = 0, pixels_cap = 0, rows_cap = 0, loss_cap = 0;
    uint32_t vtiles_cap = 0;  // tiles of the largest view (the local tile sort's key range)
    int n = 0, views = 0;
    uint64_t generation = 0;  // bumped when a buffer moves
    // the last rendered frame
    bool rendered = false;
    bool slot_f_clean = false;  // slot_f of the last frame is all zero (its key emission ran, no backward yet)
    int frame_n = 0;
    uint32_t frame_w[MAX_VIEWS] = {}, frame_h[MAX_VIEWS] = {};
    int32_t *gid_sorted = nullptr;
    // graph of the train step
    bool graph_on = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> key;
    int64_t captures = 0, replays = 0;
};

void destroy_batch_state(BatchState *b) {
    if (!b) return;
    Buf *all[] = {&b->loff, &b->block_vis, &b->dk0, &b->dk1, &b->dv0, &b->dv1, &b->ord_sums, &b->scan_part,
                  &b->splat_a, &b->splat_b, &b->splat_blue, &b->depths,   &b->rect,     &b->radii,      &b->count,
                  &b->offsets, &b->block_sums, &b->big_list, &b->big_sums, &b->keys0, &b->keys1,    &b->vals0,
                  &b->vals1,   &b->slot_a,     &b->slot_b,   &b->slot_c,   &b->slot_f, &b->hist,  &b->digit_total, &b->tile_start,
                  &b->image,   &b->final_T,    &b->final_idx, &b->v_image, &b->row_loss, &b->loss_scratch, &b->grads,
                  &b->vis_any, &b->table,      &b->scalars};
    for (Buf *x : all)
        if (x->p) (void)hipFree(x->p);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->tab_done) (void)hipEventDestroy(b->tab_done);
    if (b->h_tab) (void)hipHostFree(b->h_tab);
    if (b->h_sc) (void)hipHostFree(b->h_sc);
    delete b;
}

}  // namespace msplat

namespace {

int state(msplat_hip_ctx *ctx, BatchState *&out) {
    if (!ctx->batch) {
        BatchState *b = new BatchState();
        if (hipHostMalloc((void **)&b->h_tab, sizeof(BatchTable), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&b->h_sc, sizeof(BatchScalars), hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&b->tab_done, hipEventDisableTiming) != hipSuccess) {
            destroy_batch_state(b);
            return set_error(MSPLAT_ERR_HIP, "batch state allocation failed");
        }
        memset(b->h_tab, 0, sizeof(BatchTable));
        ctx->batch = b;
    }
    out = ctx->batch;
    return MSPLAT_OK;
}

int grow(msplat_hip_ctx *ctx, BatchState *b, Buf &buf, size_t bytes) {
    void *old = buf.p;
    CHECKED(ctx->grow(buf, bytes));
    if (buf.p != old) b->generation++;
    return MSPLAT_OK;
}

size_t record_count(BatchState *b) { return (size_t)b->views * (size_t)(b->n > 0 ? b->n : 1); }

// Radix histograms for the larger of the two sorts: the depth sort over every record
// and the tile sort over the intersection capacity.
int grow_sort_scratch(msplat_hip_ctx *ctx, BatchState *b) {
    const size_t n = std::max((size_t)b->cap, record_count(b));
    CHECKED(grow(ctx, b, b->hist, 4 * radix_hist_words((uint32_t)n)));
    CHECKED(grow(ctx, b, b->digit_total, 4 * 256));
    return MSPLAT_OK;
}

struct Layout {
    uint32_t tiles = 0, pixels = 0, rows = 0, loss_blocks = 0;
};

int check_views(const msplat_hip_gaussians *g, const msplat_hip_views *views, bool training) {
    CHECKED(check_gaussians(g));
    if (!views || !views->cameras || !views->backgrounds) return set_error(MSPLAT_ERR_ARG, "views, cameras or backgrounds is NULL");
    const int nv = views->num_views;
    if (nv < 1 || nv > MAX_VIEWS)
        return set_error(MSPLAT_ERR_ARG, "num_views must be in [1, %d], got %d", MAX_VIEWS, nv);
    if (training && !views->targets) return set_error(MSPLAT_ERR_ARG, "targets is NULL");
    for (int v = 0; v < nv; v++) {
        CHECKED(check_camera(&views->cameras[v]));
        if (views->cameras[v].degree != views->cameras[0].degree ||
            views->cameras[v].glob_scale != views->cameras[0].glob_scale)
            return set_error(MSPLAT_ERR_ARG, "the views of a batch must share degree and glob_scale");
        if (training && !views->targets[v]) return set_error(MSPLAT_ERR_ARG, "target %d is NULL", v);
    }
    if ((int64_t)nv * (int64_t)(g->num_points > 0 ? g->num_points : 1) > 0x7FFFFFFFll)
        return set_error(MSPLAT_ERR_RANGE, "views x gaussians exceeds 2^31 records");
    if (g->num_points > 0 && views->cameras[0].degree > 0 && !g->features_rest)
        return set_error(MSPLAT_ERR_ARG, "features_rest is NULL");
    return MSPLAT_OK;
}

// Fills the staging table and sizes every buffer but the intersection ones.
int prepare(msplat_hip_ctx *ctx, BatchState *b, const msplat_hip_gaussians *g, const msplat_hip_views *views,
            float ssim_weight, const msplat_hip_adam *adam) {
    const int nv = views->num_views, n = g->num_points;
    if (b->tab_pending) {  // the staging copy may still be in flight
        HIPCHECK(hipEventSynchronize(b->tab_done));
        b->tab_pending = false;
    }
    BatchTable &t = *b->h_tab;
    Layout L;
    t.num_views = nv;
    for (int v = 0; v < nv; v++) {
        const msplat_hip_camera &c = views->cameras[v];
        BatchView &bv = t.views[v];
        bv.cam = c;
        bv.gt = views->targets ? views->targets[v] : nullptr;
        for (int k = 0; k < 3; k++) bv.bg[k] = views->backgrounds[3 * v + k];
        bv.inv_n = 1.0f / (float)(3.0 * (double)c.width * (double)c.height);
        bv.inv_max_dim = 1.0f / (float)(c.width > c.height ? c.width : c.height);
        bv.pix_base = L.pixels;
        t.tile_base[v] = L.tiles;
        t.row_base[v] = L.rows;
        t.loss_base[v] = L.loss_blocks;
        const uint64_t pix = (uint64_t)L.pixels + (uint64_t)c.width * c.height;
        if (pix > 0x7FFFFFFFull) return set_error(MSPLAT_ERR_RANGE, "the batch exceeds 2^31 pixels");
        L.pixels = (uint32_t)pix;
        L.tiles += (uint32_t)(c.tiles_x * c.tiles_y);
        L.rows += c.height;
        L.loss_blocks += loss_blocks(c.width, c.height);
    }
    t.tile_base[nv] = L.tiles;
    uint32_t vt = 1;
    for (int v = 0; v < nv; v++)
        vt = std::max(vt, (uint32_t)(views->cameras[v].tiles_x * views->cameras[v].tiles_y));
    t.max_view_tiles = vt;
    t.row_base[nv] = L.rows;
    t.loss_base[nv] = L.loss_blocks;
    t.total_tiles = L.tiles;
    t.total_pixels = L.pixels;
    t.total_rows = L.rows;
    t.total_loss_blocks = L.loss_blocks;
    t.ssim_weight = ssim_weight;
    uint32_t dtu = 0;
    for (int v = 0; v < nv; v++) dtu = std::max(dtu, std::min(views->cameras[v].degrees_to_use, 3u));
    t.rest_active = 3 * (dtu >= 3 ? 15 : dtu == 2 ? 8 : dtu == 1 ? 3 : 0);
    if (adam) {
        for (int i = 0; i < 6; i++) {
            t.step[i] = adam->step_size[i];
            t.bc2[i] = adam->bc2_sqrt[i];
        }
        t.beta1 = adam->beta1;
        t.beta2 = adam->beta2;
        t.eps = adam->eps;
    }
    // buffers: per record, per tile, per pixel (grow-only, with headroom)
    const size_t R = (size_t)nv * (size_t)(n > 0 ? n : 1);
    const size_t nbx = (size_t)((n + 255) / 256);
    CHECKED(grow(ctx, b, b->splat_a, 16 * R));
    CHECKED(grow(ctx, b, b->splat_b, 16 * R));
    CHECKED(grow(ctx, b, b->splat_blue, 4 * R));
    CHECKED(grow(ctx, b, b->depths, 4 * R));
    CHECKED(grow(ctx, b, b->rect, 8 * R));
    CHECKED(grow(ctx, b, b->radii, 4 * R));
    CHECKED(grow(ctx, b, b->count, 4 * R));
    CHECKED(grow(ctx, b, b->offsets, 4 * R));
    CHECKED(grow(ctx, b, b->block_sums, 4 * (nbx * nv + 1)));
    CHECKED(grow(ctx, b, b->block_vis, 4 * (nbx * nv + 1)));
    CHECKED(grow(ctx, b, b->dk0, 4 * R));
    CHECKED(grow(ctx, b, b->dk1, 4 * R));
    CHECKED(grow(ctx, b, b->dv0, 4 * R));
    CHECKED(grow(ctx, b, b->dv1, 4 * R));
    CHECKED(grow(ctx, b, b->loff, 4 * R));
    CHECKED(grow(ctx, b, b->ord_sums, 4 * ((R + 255) / 256 + 1)));
    CHECKED(grow(ctx, b, b->scan_part, 4 * 3 * scan_partial_words((uint32_t)std::max(nbx * nv, (R + 255) / 256)) + 64));
    CHECKED(grow(ctx, b, b->big_list, 4 * R));
    CHECKED(grow(ctx, b, b->big_sums, 36 * R + 64));
    b->tiles_cap = std::max(b->tiles_cap, std::max(L.tiles, vt * (uint32_t)nv));  // items and ranges
    b->vtiles_cap = std::max(b->vtiles_cap, vt);
    b->pixels_cap = std::max(b->pixels_cap, L.pixels);
    b->rows_cap = std::max(b->rows_cap, L.rows);
    b->loss_cap = std::max(b->loss_cap, L.loss_blocks);
    CHECKED(grow(ctx, b, b->tile_start, 4 * ((size_t)b->tiles_cap + 1)));
    const size_t P = b->pixels_cap > 0 ? b->pixels_cap : 1;
    CHECKED(grow(ctx, b, b->image, 12 * P));
    CHECKED(grow(ctx, b, b->final_T, 4 * P));
    CHECKED(grow(ctx, b, b->final_idx, 4 * P));
    CHECKED(grow(ctx, b, b->v_image, 12 * P));
    CHECKED(grow(ctx, b, b->row_loss, 24 * ((size_t)b->rows_cap + 1)));
    if (adam) CHECKED(grow(ctx, b, b->loss_scratch, 4 * (size_t)LOSS_PLANES * P));  // training only
    CHECKED(grow(ctx, b, b->table, sizeof(BatchTable)));
    CHECKED(grow(ctx, b, b->scalars, sizeof(BatchScalars)));
    b->n = n;
    b->views = nv;
    CHECKED(grow_sort_scratch(ctx, b));  // the record count may have grown
    (void)batch_raster_grid(false, ctx->tuning);  // occupancy queries outside any capture
    (void)batch_raster_grid(true, ctx->tuning);
    HIPCHECK(hipMemcpyAsync(b->table.p, b->h_tab, sizeof(BatchTable), hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipEventRecord(b->tab_done, ctx->stream));
    b->tab_pending = true;
    return MSPLAT_OK;
}

int grow_isects(msplat_hip_ctx *ctx, BatchState *b, uint64_t need) {
    const uint64_t want = need + need / 4 + 4096;
    if (want > 0x7FFFFFFFull)
        return set_error(MSPLAT_ERR_RANGE, "more than 2^31 tile intersections (%llu); lower the resolution",
                         (unsigned long long)need);
    if (want <= b->cap && b->cap) return MSPLAT_OK;
    const size_t C = (size_t)want;
    CHECKED(grow(ctx, b, b->keys0, 4 * C));
    CHECKED(grow(ctx, b, b->keys1, 4 * C));
    CHECKED(grow(ctx, b, b->vals0, 4 * C));
    CHECKED(grow(ctx, b, b->vals1, 4 * C));
    CHECKED(grow(ctx, b, b->slot_a, 16 * C));
    CHECKED(grow(ctx, b, b->slot_b, 16 * C));
    CHECKED(grow(ctx, b, b->slot_c, 4 * C));
    CHECKED(grow(ctx, b, b->slot_f, (C + 3) & ~(uint64_t)3));
    b->cap = (uint32_t)C;
    return grow_sort_scratch(ctx, b);
}

BatchScalars *dsc(BatchState *b) { return (BatchScalars *)b->scalars.p; }
const BatchTable *dtab(BatchState *b) { return (const BatchTable *)b->table.p; }

BatchProjectArgs project_args(BatchState *b, const msplat_hip_gaussians *g) {
    BatchProjectArgs a;
    a.tab = dtab(b);
    a.n = g->num_points;
    a.num_views = b->views;
    a.nbx = (g->num_points + 255) / 256;
    a.means = g->means;
    a.scales = g->scales;
    a.quats = g->quats;
    a.dc = g->features_dc;
    a.rest = g->features_rest;
    a.opac = g->opacities;
    a.splat_a = (float4 *)b->splat_a.p;
    a.splat_b = (float4 *)b->splat_b.p;
    a.splat_blue = (float *)b->splat_blue.p;
    a.depths = (float *)b->depths.p;
    a.rect = (uint2 *)b->rect.p;
    a.radii = (int32_t *)b->radii.p;
    a.count = (uint32_t *)b->count.p;
    a.block_sums = (uint32_t *)b->block_sums.p;
    a.block_vis = (uint32_t *)b->block_vis.p;
    uint32_t dtu = 0, deg = 0;
    for (int v = 0; v < b->views; v++) {
        dtu = std::max(dtu, b->h_tab->views[v].cam.degrees_to_use);
        deg = std::max(deg, b->h_tab->views[v].cam.degree);
    }
    a.sh_u = deg > 3 || dtu == 0 ? 0 : 3 * (int)sh_bases(dtu);  // degree 4 or the DC alone: read in place
    a.sh_stride = 3 * (size_t)(sh_bases(deg) - 1);
    return a;
}

BatchRasterArgs raster_args(BatchState *b) {
    BatchRasterArgs r;
    r.tab = dtab(b);
    r.num_views = b->views;
    r.queue = &dsc(b)->queue[0][0];
    r.tile_start = (const int32_t *)b->tile_start.p;
    r.gid = b->gid_sorted;
    r.splat_a = (const float4 *)b->splat_a.p;
    r.splat_b = (const float4 *)b->splat_b.p;
    r.splat_blue = (const float *)b->splat_blue.p;
    r.rect = (const uint2 *)b->rect.p;
    r.offsets = (const uint32_t *)b->offsets.p;
    r.image = (float *)b->image.p;
    r.final_T = (float *)b->final_T.p;
    r.final_idx = (int32_t *)b->final_idx.p;
    r.v_image = (const float *)b->v_image.p;
    r.slot_a = (float4 *)b->slot_a.p;
    r.slot_b = (float4 *)b->slot_b.p;
    r.slot_c = (float *)b->slot_c.p;
    r.slot_f = (uint8_t *)b->slot_f.p;
    return r;
}

float4 *big_a(BatchState *b) { return (float4 *)(((uintptr_t)b->big_sums.p + 15) & ~(uintptr_t)15); }

uint32_t *scan_parts(BatchState *b, int i) {
    const size_t pw = scan_partial_words((uint32_t)std::max((size_t)b->views * ((b->n + 255) / 256),
                                                            (record_count(b) + 255) / 256));
    return (uint32_t *)b->scan_part.p + i * pw;
}

// projection and the count scans (intersections against the capacity, visible records)
void enqueue_counts(msplat_hip_ctx *ctx, BatchState *b, const msplat_hip_gaussians *g, uint32_t cap) {
    hipStream_t s = ctx->stream;
    launch_batch_project(project_args(b, g), s);
    const uint32_t L = (uint32_t)b->views * (uint32_t)((g->num_points + 255) / 256);
    BatchScalars *sc = dsc(b);
    launch_scan_u32((uint32_t *)b->block_sums.p, L, scan_parts(b, 0),
                    ScanOut{&sc->total, &sc->total32, cap, &sc->overflow, &sc->big_count}, s);
    launch_scan_u32((uint32_t *)b->block_vis.p, L, scan_parts(b, 1),
                    ScanOut{&sc->nvis64, &sc->nvis, 0xFFFFFFFFu, nullptr, nullptr}, s);
}

// visible records sorted by depth, keys in depth order, tile sort, ranges, forward
void enqueue_sort_and_forward(msplat_hip_ctx *ctx, BatchState *b, const msplat_hip_gaussians *g) {
    hipStream_t s = ctx->stream;
    BatchScalars *sc = dsc(b);
    const uint32_t R = (uint32_t)record_count(b);
    CompactArgs c;
    c.n = g->num_points;
    c.num_views = b->views;
    c.nbx = (g->num_points + 255) / 256;
    c.count = (const uint32_t *)b->count.p;
    c.depths = (const float *)b->depths.p;
    c.block_vis = (const uint32_t *)b->block_vis.p;
    c.dkeys = (uint32_t *)b->dk0.p;
    c.dvals = (int32_t *)b->dv0.p;
    launch_compact(c, s);
    uint32_t *dki = (uint32_t *)b->dk0.p, *dko = (uint32_t *)b->dk1.p;
    int32_t *dvi = (int32_t *)b->dv0.p, *dvo = (int32_t *)b->dv1.p;
    for (int shift = 0; shift < 32; shift += 8) {
        launch_radix_pass(dki, dvi, dko, dvo, &sc->nvis, R, shift, (uint32_t *)b->hist.p,
                          (uint32_t *)b->digit_total.p, s);
        std::swap(dki, dko);
        std::swap(dvi, dvo);
    }
    // stably by view: each view's records in depth order, views in order
    launch_view_keys(dvi, &sc->nvis, R, g->num_points, dki, s);
    launch_radix_pass(dki, dvi, dko, dvo, &sc->nvis, R, 0, (uint32_t *)b->hist.p, (uint32_t *)b->digit_total.p, s);
    std::swap(dki, dko);
    std::swap(dvi, dvo);
    launch_order_sums(dvi, &sc->nvis, R, (const uint32_t *)b->count.p, (uint32_t *)b->ord_sums.p,
                      (uint32_t *)b->loff.p, s);
    launch_scan_u32((uint32_t *)b->ord_sums.p, (R + 255) / 256, scan_parts(b, 2),
                    ScanOut{nullptr, &sc->total32_b, b->cap, &sc->overflow_b, nullptr}, s);
    KeygenArgs k;
    memset(&k, 0, sizeof k);
    k.list = dvi;
    k.nvis_ptr = &sc->nvis;
    k.nvis_max = R;
    k.count = (const uint32_t *)b->count.p;
    k.rect = (const uint2 *)b->rect.p;
    k.block_bases = (const uint32_t *)b->ord_sums.p;
    k.offsets = (uint32_t *)b->offsets.p;
    k.loff = (uint32_t *)b->loff.p;
    k.m_ptr = &sc->total32;
    k.m_max = b->cap;
    k.keys = (uint32_t *)b->keys0.p;
    k.vals = (int32_t *)b->vals0.p;
    k.big_list = (uint32_t *)b->big_list.p;
    k.big_count = &sc->big_count;
    k.overflow = &sc->overflow;
    k.tab = dtab(b);
    k.n = g->num_points;
    k.slot_f = (uint8_t *)b->slot_f.p;
    launch_keygen(k, true, s);
    uint32_t *kin = (uint32_t *)b->keys0.p, *kout = (uint32_t *)b->keys1.p;
    int32_t *vin = (int32_t *)b->vals0.p, *vout = (int32_t *)b->vals1.p;
    const int bits = tile_bits((int)b->vtiles_cap);
    for (int done = 0; done < bits;) {  // narrow first pass, as for one view (abi.hip)
        const int width = done == 0 && bits % 8 ? bits % 8 : 8;
        launch_radix_pass(kin, vin, kout, vout, &sc->total32, b->cap, done, (uint32_t *)b->hist.p,
                          (uint32_t *)b->digit_total.p, s, (1u << width) - 1u);
        std::swap(kin, kout);
        std::swap(vin, vout);
        done += width;
    }
    b->gid_sorted = vin;
    launch_tile_ranges_views(kin, vin, &sc->total32, b->cap, g->num_points, dtab(b), (int32_t *)b->tile_start.p, s);
    launch_batch_raster_forward(raster_args(b), ctx->tuning, s);
}

// raster backward, big-list sums, per-gaussian sums over the views. The slots' zero flags
// are all zero after the frame's key emission; a later backward of the same frame
// (clear_flags) zeroes them first.
void enqueue_backward(msplat_hip_ctx *ctx, BatchState *b, const msplat_hip_gaussians *g, float *const gout[6],
                      const msplat_hip_grad_stats *stats, bool clear_flags) {
    hipStream_t s = ctx->stream;
    if (clear_flags) {
        const size_t words = (b->cap + 3) / 4;
        if (words) zero_u32<<<(unsigned)std::min<size_t>((words + 255) / 256, 2048), 256, 0, s>>>((uint32_t *)b->slot_f.p, words);
    }
    launch_batch_raster_backward(raster_args(b), ctx->tuning, s);
    const size_t R = record_count(b);
    float4 *ba = big_a(b);
    launch_big_reduce_records((const uint32_t *)b->big_list.p, &dsc(b)->big_count, (const uint32_t *)b->offsets.p,
                              (const uint32_t *)b->count.p, (const float4 *)b->slot_a.p, (const float4 *)b->slot_b.p,
                              (const float *)b->slot_c.p, (const uint8_t *)b->slot_f.p, ba, ba + R, (float *)(ba + 2 * R), s);
    BatchBackwardArgs a;
    a.tab = dtab(b);
    a.n = g->num_points;
    a.num_views = b->views;
    a.means = g->means;
    a.scales = g->scales;
    a.quats = g->quats;
    a.radii = (const int32_t *)b->radii.p;
    a.count = (const uint32_t *)b->count.p;
    a.offsets = (const uint32_t *)b->offsets.p;
    a.splat_a = (const float4 *)b->splat_a.p;
    a.splat_b = (const float4 *)b->splat_b.p;
    a.slot_a = (const float4 *)b->slot_a.p;
    a.slot_b = (const float4 *)b->slot_b.p;
    a.slot_c = (const float *)b->slot_c.p;
    a.slot_f = (const uint8_t *)b->slot_f.p;
    a.big_a = ba;
    a.big_b = ba + R;
    a.big_c = (const float *)(ba + 2 * R);
    a.sc = dsc(b);
    a.rest_stride = (sh_bases(b->h_tab->views[0].cam.degree) - 1) * 3;
    a.g_means = gout[0];
    a.g_scales = gout[1];
    a.g_quats = gout[2];
    a.g_dc = gout[3];
    a.g_rest = gout[4];
    a.g_opac = gout[5];
    a.vis_any = (uint8_t *)b->vis_any.p;
    a.vis = stats ? stats->vis_counts : nullptr;
    a.gn = stats ? stats->xys_grad_norm : nullptr;
    a.ms = stats ? stats->max_2d_size : nullptr;
    launch_batch_backward(a, s);
}

// scratch gradient buffers of the six groups in b->grads
int scratch_grads(msplat_hip_ctx *ctx, BatchState *b, int n, uint32_t rest_stride, float *out[6]) {
    const size_t N = (size_t)(n > 0 ? n : 1);
    const size_t w[6] = {3, 3, 4, 3, rest_stride, 1};
    size_t tot = 0;
    for (int i = 0; i < 6; i++) tot += w[i] * N;
    CHECKED(grow(ctx, b, b->grads, 4 * tot + 64));
    CHECKED(grow(ctx, b, b->vis_any, N + 64));
    float *p = (float *)b->grads.p;
    for (int i = 0; i < 6; i++) {
        out[i] = p;
        p += w[i] * N;
    }
    return MSPLAT_OK;
}

// The step after the count pass (counts: also enqueue it, against the capacity)
void enqueue_train(msplat_hip_ctx *ctx, BatchState *b, const msplat_hip_gaussians *g, const msplat_hip_adam *adam,
                   const msplat_hip_grad_stats *stats, float *const gr[6], bool counts) {
    hipStream_t s = ctx->stream;
    if (counts) enqueue_counts(ctx, b, g, b->cap);
    enqueue_sort_and_forward(ctx, b, g);
    BatchLossArgs l;
    l.tab = dtab(b);
    l.num_views = b->views;
    l.max_loss_blocks = b->loss_cap;
    l.max_rows = b->rows_cap;
    l.img = (const float *)b->image.p;
    l.v_img = (float *)b->v_image.p;
    l.scratch = (float *)b->loss_scratch.p;
    l.row_loss = (double *)b->row_loss.p;
    l.sc = dsc(b);
    launch_batch_loss(l, s);
    enqueue_backward(ctx, b, g, gr, stats, false);
    const size_t N = (size_t)g->num_points;
    const uint32_t stride = (sh_bases(b->h_tab->views[0].cam.degree) - 1) * 3;
    float *params[6] = {g->means, g->scales, g->quats, g->features_dc, g->features_rest, g->opacities};
    const int width[6] = {3, 3, 4, 3, (int)stride, 1};
    const uint8_t *vis = (const uint8_t *)b->vis_any.p;
    for (int i = 0; i < 6; i++) {
        if (width[i] == 0) continue;
        launch_batch_adam(params[i], adam->exp_avg[i], adam->exp_avg_sq[i], gr[i], N * (size_t)width[i], width[i],
                          (i == 3 || i == 4) ? vis : nullptr, dtab(b), i, &dsc(b)->overflow, s);
    }
}

// Reads the frame's scalars back (syncs); *over = the capacity was exceeded.
int read_scalars(msplat_hip_ctx *ctx, BatchState *b, bool &over) {
    HIPCHECK(hipMemcpyAsync(b->h_sc, b->scalars.p, sizeof(BatchScalars), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    b->tab_pending = false;
    over = b->h_sc->overflow != 0;
    return MSPLAT_OK;
}

// Sizes the intersection buffers from a counting pass when they are not sized yet
int ensure_capacity(msplat_hip_ctx *ctx, BatchState *b, const msplat_hip_gaussians *g) {
    if (b->cap) return MSPLAT_OK;
    enqueue_counts(ctx, b, g, 0u);  // capacity 0: counts only
    HIPCHECK(hipGetLastError());
    bool over = false;
    CHECKED(read_scalars(ctx, b, over));
    return grow_isects(ctx, b, b->h_sc->total);
}

// Eager frames count first and read the totals back (one sync, as a one-view frame
// does), so their buffers always fit and a frame never runs twice.
int eager_counts(msplat_hip_ctx *ctx, BatchState *b, const msplat_hip_gaussians *g) {
    enqueue_counts(ctx, b, g, 0xFFFFFFFFu);
    HIPCHECK(hipGetLastError());
    bool over = false;
    CHECKED(read_scalars(ctx, b, over));
    return grow_isects(ctx, b, b->h_sc->total);
}

int validate_adam(const msplat_hip_gaussians *g, const msplat_hip_adam *adam, uint32_t stride) {
    for (int i = 0; i < 6; i++) {
        if (i == 4 && stride == 0) continue;
        if (g->num_points > 0 && (!adam->exp_avg[i] || !adam->exp_avg_sq[i]))
            return set_error(MSPLAT_ERR_ARG, "Adam moment %d is NULL", i);
    }
    return MSPLAT_OK;
}

std::vector<uint64_t> graph_key(msplat_hip_ctx *ctx, BatchState *b, const msplat_hip_gaussians *g,
                                const msplat_hip_adam *adam, const msplat_hip_grad_stats *st) {
    std::vector<uint64_t> k = {(uint64_t)b->n,          (uint64_t)b->views,       b->cap,
                               b->tiles_cap,            b->pixels_cap,            b->rows_cap,
                               b->loss_cap,             b->generation,            b->vtiles_cap,            (uint64_t)ctx->tuning.bwd_ppt,
                               (uint64_t)ctx->tuning.fwd_batch, (uint64_t)ctx->tuning.bwd_batch,
                               (uint64_t)ctx->tuning.cull,
                               b->h_tab->views[0].cam.degree};
    uint32_t dtu = 0;  // the SH floats the projection stages (project_args)
    for (int v = 0; v < b->views; v++) dtu = std::max(dtu, b->h_tab->views[v].cam.degrees_to_use);
    k.push_back(dtu);
    const void *ptrs[] = {g->means, g->scales, g->quats, g->features_dc, g->features_rest, g->opacities,
                          st->vis_counts, st->xys_grad_norm, st->max_2d_size};
    for (const void *p : ptrs) k.push_back((uint64_t)(uintptr_t)p);
    for (int i = 0; i < 6; i++) {
        k.push_back((uint64_t)(uintptr_t)adam->exp_avg[i]);
        k.push_back((uint64_t)(uintptr_t)adam->exp_avg_sq[i]);
    }
    return k;
}

int run_train(msplat_hip_ctx *ctx, BatchState *b, const msplat_hip_gaussians *g, const msplat_hip_adam *adam,
              const msplat_hip_grad_stats *stats) {
    const uint32_t stride = (sh_bases(b->h_tab->views[0].cam.degree) - 1) * 3;
    float *gr[6];
    CHECKED(scratch_grads(ctx, b, g->num_points, stride, gr));
    if (!b->graph_on) {
        CHECKED(eager_counts(ctx, b, g));
        enqueue_train(ctx, b, g, adam, stats, gr, false);
        HIPCHECK(hipGetLastError());
        return MSPLAT_OK;
    }
    std::vector<uint64_t> key = graph_key(ctx, b, g, adam, stats);
    if (!b->exec || key != b->key) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        HIPCHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        enqueue_train(ctx, b, g, adam, stats, gr, true);
        const hipError_t launch_err = hipGetLastError();
        hipGraph_t graph = nullptr;
        const hipError_t end_err = hipStreamEndCapture(ctx->stream, &graph);
        if (launch_err != hipSuccess || end_err != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return set_error(MSPLAT_ERR_HIP, "graph capture failed: %s",
                             hipGetErrorString(launch_err != hipSuccess ? launch_err : end_err));
        }
        b->graph = graph;
        HIPCHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->key = key;
        b->captures++;
    }
    HIPCHECK(hipGraphLaunch(b->exec, ctx->stream));
    b->replays++;
    return MSPLAT_OK;
}

}  // namespace

extern "C" {

int msplat_hip_render_batch(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_views *views,
                            float **out_images, uint64_t *pixel_offsets) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_replays(ctx));
    CHECKED(check_views(g, views, false));
    BatchState *b;
    CHECKED(state(ctx, b));
    b->rendered = false;
    CHECKED(prepare(ctx, b, g, views, 0.f, nullptr));
    CHECKED(eager_counts(ctx, b, g));
    enqueue_sort_and_forward(ctx, b, g);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    b->rendered = true;
    b->slot_f_clean = true;
    b->frame_n = g->num_points;
    for (int v = 0; v < views->num_views; v++) {
        b->frame_w[v] = views->cameras[v].width;
        b->frame_h[v] = views->cameras[v].height;
        if (pixel_offsets) pixel_offsets[v] = b->h_tab->views[v].pix_base;
    }
    if (out_images) *out_images = (float *)b->image.p;
    return MSPLAT_OK;
}

int msplat_hip_train_step_batch(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_views *views,
                                float ssim_weight, const msplat_hip_adam *adam, const msplat_hip_grad_stats *stats,
                                float *losses) {
    if (!ctx || !adam || !stats) return set_error(MSPLAT_ERR_ARG, "ctx, adam or stats is NULL");
    CHECKED(settle_replays(ctx));
    CHECKED(check_views(g, views, true));
    BatchState *b;
    CHECKED(state(ctx, b));
    if (b->graph_on && !ctx->stream) return set_error(MSPLAT_ERR_ARG, "graph capture needs a non-null stream");
    b->rendered = false;
    CHECKED(prepare(ctx, b, g, views, ssim_weight, adam));
    CHECKED(validate_adam(g, adam, (sh_bases(views->cameras[0].degree) - 1) * 3));
    if (b->graph_on) CHECKED(ensure_capacity(ctx, b, g));
    for (int attempt = 0; attempt < 2; attempt++) {
        CHECKED(run_train(ctx, b, g, adam, stats));
        bool over = false;
        CHECKED(read_scalars(ctx, b, over));
        if (!over) break;  // over capacity the step changed nothing: grow and run it again
        if (attempt == 1) return set_error(MSPLAT_ERR_HIP, "intersection capacity still exceeded after growth");
        CHECKED(grow_isects(ctx, b, b->h_sc->total));
    }
    if (losses)
        for (int v = 0; v < views->num_views; v++) losses[v] = b->h_sc->loss_mean[v];
    return MSPLAT_OK;
}

int msplat_hip_render_backward_batch(msplat_hip_ctx *ctx, const msplat_hip_gaussians *g, const msplat_hip_views *views,
                                     const float *v_images, const msplat_hip_grads *out) {
    if (!ctx || !v_images || !out) return set_error(MSPLAT_ERR_ARG, "ctx, v_images or out is NULL");
    CHECKED(settle_replays(ctx));
    CHECKED(check_views(g, views, false));
    BatchState *b = ctx->batch;
    if (!b || !b->rendered || b->frame_n != g->num_points || b->views != views->num_views)
        return set_error(MSPLAT_ERR_ARG, "render_backward_batch does not match the last render_batch on this context");
    for (int v = 0; v < views->num_views; v++)
        if (b->frame_w[v] != views->cameras[v].width || b->frame_h[v] != views->cameras[v].height)
            return set_error(MSPLAT_ERR_ARG, "render_backward_batch does not match the last render_batch (view %d)", v);
    HIPCHECK(hipMemcpyAsync(b->v_image.p, v_images, 12 * (size_t)b->h_tab->total_pixels, hipMemcpyDeviceToDevice,
                            ctx->stream));
    const uint32_t stride = (sh_bases(views->cameras[0].degree) - 1) * 3;
    float *scratch[6];
    CHECKED(scratch_grads(ctx, b, g->num_points, stride, scratch));
    float *user[6] = {out->v_means, out->v_scales, out->v_quats, out->v_features_dc, out->v_features_rest,
                      out->v_opacities};
    float *gr[6];
    for (int i = 0; i < 6; i++) gr[i] = user[i] ? user[i] : scratch[i];
    enqueue_backward(ctx, b, g, gr, nullptr, !b->slot_f_clean);
    b->slot_f_clean = false;
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    return MSPLAT_OK;
}

int msplat_hip_set_batch_graph(msplat_hip_ctx *ctx, int enable) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    BatchState *b;
    CHECKED(state(ctx, b));
    b->graph_on = enable != 0;
    return MSPLAT_OK;
}

int msplat_hip_batch_graph_stats(msplat_hip_ctx *ctx, int64_t *captures, int64_t *replays) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    BatchState *b = ctx->batch;
    if (captures) *captures = b ? b->captures : 0;
    if (replays) *replays = b ? b->replays : 0;
    return MSPLAT_OK;
}

int msplat_hip_batch_frame_info(msplat_hip_ctx *ctx, uint64_t *num_isects, uint64_t *num_visible) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_replays(ctx));
    BatchState *b = ctx->batch;
    if (num_isects) *num_isects = b ? b->h_sc->total : 0;
    if (num_visible) *num_visible = b ? b->h_sc->nvis64 : 0;
    return MSPLAT_OK;
}

}  // extern "C"
