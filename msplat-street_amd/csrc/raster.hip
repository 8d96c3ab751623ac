// Rasterizer kernels: one workgroup per tile for a single view, and persistent
// workgroups that pull (view, tile) items from a device work queue for a batch of views
// of mixed resolution. Both run the tile bodies of raster_tile.h.
#include "raster_tile.h"

namespace msplat {

namespace {

__device__ inline TileData tile_data(const RasterArgs &a) {
    return TileData{a.tile_start, a.gid,    a.gid_mask, a.splat_a, a.splat_b, a.splat_blue,     a.rect,
                    a.offsets,    a.slot_a, a.slot_b,   a.slot_c,  a.slot_f,  a.slots_written, a.cull != 0,
                    a.depths,     a.slot_d};
}

__device__ inline TileView tile_view(const RasterArgs &a) {
    return TileView{a.width,   a.height, a.tiles_x, {a.bg[0], a.bg[1], a.bg[2]}, a.image, a.final_T, a.final_idx,
                    a.v_image, a.bg_img, a.out_depth, a.v_aux};
}

__device__ inline TileData tile_data(const BatchRasterArgs &a) {
    return TileData{a.tile_start, a.gid,    0xFFFFFFFFu, a.splat_a, a.splat_b, a.splat_blue, a.rect,
                    a.offsets,    a.slot_a, a.slot_b,    a.slot_c,  a.slot_f,  nullptr,      a.cull != 0};
}

// View v of a batch: per-pixel buffers offset to the view's first pixel
__device__ inline TileView tile_view(const BatchRasterArgs &a, int v) {
    const BatchView &bv = a.tab->views[v];
    const size_t pb = bv.pix_base;
    return TileView{bv.cam.width,
                    bv.cam.height,
                    bv.cam.tiles_x,
                    {bv.bg[0], bv.bg[1], bv.bg[2]},
                    a.image ? a.image + 3 * pb : nullptr,
                    a.final_T + pb,
                    a.final_idx + pb,
                    a.v_image ? a.v_image + 3 * pb : nullptr};
}

// The view of global tile t: the last v with tile_base[v] <= t
__device__ inline int view_of(const uint32_t *tb, int nv, uint32_t t) {
    int lo = 0, hi = nv - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tb[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Work queue of a persistent launch: the items [0, total) are cut into NUM_XCD
// contiguous chunks with one head each. A workgroup drains the chunk of its own XCD
// first (neighbouring tiles then share an L2) and then helps with the others; it
// returns total once every chunk is drained. One atomic per item.
struct ItemQueue {
    uint32_t *heads;  // NUM_XCD counters, zeroed before the launch
    uint32_t total;
    int chunk;        // the chunk being drained, NUM_XCD when done
    int home;
};

__device__ inline uint32_t next_item(ItemQueue &q, uint32_t *slot) {
    for (;;) {
        if (q.chunk >= NUM_XCD) return q.total;
        const int c = (q.home + q.chunk) % NUM_XCD;
        const uint32_t per = q.total / NUM_XCD, rem = q.total % NUM_XCD;
        const uint32_t len = per + ((uint32_t)c < rem ? 1u : 0u);
        const uint32_t first = (uint32_t)c * per + ((uint32_t)c < rem ? (uint32_t)c : rem);
        if (threadIdx.x == 0) *slot = atomicAdd(&q.heads[c], 1u);
        __syncthreads();
        const uint32_t i = *slot;
        __syncthreads();
        if (i < len) return first + i;
        q.chunk++;
    }
}

}  // namespace

// Forward kernels are capped at 64 VGPRs (8 waves per SIMD): the unrolled compositing
// loop would otherwise hold more registers than the extra waves are worth.
template <int FB, bool AUX = false>
__global__ __launch_bounds__(256, 8) void raster_forward_kernel(RasterArgs a) {
    const int t = xcd_item(blockIdx.x, gridDim.x);
    forward_tile<FB, AUX>(tile_view(a), tile_data(a), t, t);
}

template <int PPT, int BB, bool AUX = false>
__global__ __launch_bounds__(256 / PPT) void raster_backward_kernel(RasterArgs a) {
    const int t = xcd_item(blockIdx.x, gridDim.x);
    backward_tile<PPT, BB, false, AUX>(tile_view(a), tile_data(a), t, t);
}

__device__ inline ChunkView chunk_view(const RasterArgs &a, int k) {
    return ChunkView{k, a.chunks, (size_t)a.width * a.height, a.chunk_prod, a.chunk_state, a.chunk_last};
}

// Depth-chunked kernels: workgroup b runs chunk b / tiles of tile b % tiles (tiles in
// XCD-chunked order within a chunk).
template <int FB, bool PRODUCT>
__global__ __launch_bounds__(256, 8) void raster_forward_chunk_kernel(RasterArgs a) {
    const int nt = a.tiles_x * a.tiles_y;
    const int k = blockIdx.x / nt, t = xcd_item(blockIdx.x % nt, nt);
    const ChunkView ch = chunk_view(a, k);
    forward_chunk<FB, PRODUCT>(tile_view(a), tile_data(a), ch, t, t);
}

// The per-pixel chunk passes below walk a pixel's K chunk planes in groups of CG: a
// group's loads are issued together, then used in chunk order, so a pixel waits for
// memory K / CG times instead of K times (with few tiles there are few pixels to hide
// the latency behind).
constexpr int CG = 8;

// The merge of the chunked forward, one thread per pixel: colours summed in chunk order
// up to the chunk the pixel terminated in, the final transmittance and last index, the
// composite over the background.
__global__ __launch_bounds__(256) void raster_chunk_merge_kernel(RasterArgs a) {
    const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= a.width || y >= a.height) return;
    const size_t pid = (size_t)y * a.width + x, np = (size_t)a.width * a.height;
    const int tile = (int)(y / TILE) * a.tiles_x + (int)(x / TILE);
    float cr = 0.f, cg = 0.f, cb = 0.f, T = 1.f;
    int last = a.tile_start[tile] - 1;
    bool done = false;
    for (int k0 = 0; k0 < a.chunks && !done; k0 += CG) {
        float4 c[CG];
        int l[CG];
#pragma unroll
        for (int u = 0; u < CG; u++) {
            const size_t o = (size_t)min(k0 + u, a.chunks - 1) * np + pid;
            c[u] = a.chunk_state[o];
            l[u] = a.chunk_last[o];
        }
#pragma unroll
        for (int u = 0; u < CG; u++) {
            if (done || k0 + u >= a.chunks) break;
            cr += c[u].x;
            cg += c[u].y;
            cb += c[u].z;
            T = fabsf(c[u].w);
            last = max(last, l[u]);
            done = c[u].w < 0.f;  // terminated in this chunk
        }
    }
    a.final_T[pid] = T;
    a.final_idx[pid] = last;
    a.image[3 * pid] = fminf(fmaxf(a.bg[0] * T + cr, 0.f), 1.f);
    a.image[3 * pid + 1] = fminf(fmaxf(a.bg[1] * T + cg, 0.f), 1.f);
    a.image[3 * pid + 2] = fminf(fmaxf(a.bg[2] * T + cb, 0.f), 1.f);
}

// Between the product and composite passes, per pixel: each chunk's product replaced by
// the transmittance at the chunk's start, the product of the earlier chunks' in order.
__global__ __launch_bounds__(256) void raster_chunk_prefix_kernel(RasterArgs a) {
    const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= a.width || y >= a.height) return;
    const size_t pid = (size_t)y * a.width + x, np = (size_t)a.width * a.height;
    float T = 1.f;
    for (int k0 = 0; k0 < a.chunks; k0 += CG) {
        float p[CG];
#pragma unroll
        for (int u = 0; u < CG; u++) p[u] = a.chunk_prod[(size_t)min(k0 + u, a.chunks - 1) * np + pid];
#pragma unroll
        for (int u = 0; u < CG; u++) {
            if (k0 + u >= a.chunks) break;
            a.chunk_prod[(size_t)(k0 + u) * np + pid] = T;
            T *= p[u];
        }
    }
}

// Before the chunked backward, per pixel: S at each chunk's end,
//   S_k = T_final <bg, v> + sum over chunks k < j <= terminating chunk of <colour_j, v>,
// summed back to front into the product planes. A chunk with a negated end transmittance
// terminated the pixel: what lies behind it is not in the image, so S restarts there
// (chunks behind the terminating one composite nothing and their S is never used).
__global__ __launch_bounds__(256) void raster_chunk_suffix_kernel(RasterArgs a) {
    const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= a.width || y >= a.height) return;
    const size_t pid = (size_t)y * a.width + x, np = (size_t)a.width * a.height;
    const float vr = a.v_image[3 * pid], vg = a.v_image[3 * pid + 1], vb = a.v_image[3 * pid + 2];
    const float base = a.final_T[pid] * (a.bg[0] * vr + a.bg[1] * vg + a.bg[2] * vb);
    float S = base;
    for (int k0 = a.chunks - 1; k0 >= 0; k0 -= CG) {
        float4 c[CG];
#pragma unroll
        for (int u = 0; u < CG; u++) c[u] = a.chunk_state[(size_t)max(k0 - u, 0) * np + pid];
#pragma unroll
        for (int u = 0; u < CG; u++) {
            if (k0 - u < 0) break;
            if (c[u].w < 0.f) S = base;
            a.chunk_prod[(size_t)(k0 - u) * np + pid] = S;
            S += c[u].x * vr + c[u].y * vg + c[u].z * vb;
        }
    }
}

template <int PPT, int BB>
__global__ __launch_bounds__(256 / PPT) void raster_backward_chunk_kernel(RasterArgs a) {
    const int nt = a.tiles_x * a.tiles_y;
    const int k = blockIdx.x / nt, t = xcd_item(blockIdx.x % nt, nt);
    const ChunkView ch = chunk_view(a, k);
    backward_tile<PPT, BB, true>(tile_view(a), tile_data(a), t, t, &ch);
}

// Persistent forward over every (view, tile) of a batch; item t is tile t - tile_base[v]
// of view v, whose range is number (tile * num_views + v). Every workgroup loops until
// the queue is drained, so the grid drains on its own.
template <int FB>
__global__ __launch_bounds__(256, 8) void raster_forward_batch_kernel(BatchRasterArgs a) {
    __shared__ uint32_t tb[MAX_VIEWS + 1];
    __shared__ uint32_t item_slot;
    const int nv = a.num_views;
    for (int i = threadIdx.x; i <= nv; i += 256) tb[i] = a.tab->tile_base[i];
    const uint32_t total = a.tab->total_tiles;
    const TileData d = tile_data(a);
    ItemQueue q{a.queue, total, 0, (int)(blockIdx.x % NUM_XCD)};
    for (;;) {
        const uint32_t t = next_item(q, &item_slot);
        if (t >= total) break;
        const int v = view_of(tb, nv, t);
        const int lt = (int)(t - tb[v]);
        forward_tile<FB>(tile_view(a, v), d, lt * nv + v, lt);
    }
}

template <int PPT, int BB>
__global__ __launch_bounds__(256 / PPT) void raster_backward_batch_kernel(BatchRasterArgs a) {
    __shared__ uint32_t tb[MAX_VIEWS + 1];
    __shared__ uint32_t item_slot;
    const int nv = a.num_views;
    for (int i = threadIdx.x; i <= nv; i += 256 / PPT) tb[i] = a.tab->tile_base[i];
    const uint32_t total = a.tab->total_tiles;
    const TileData d = tile_data(a);
    ItemQueue q{a.queue + NUM_XCD, total, 0, (int)(blockIdx.x % NUM_XCD)};
    for (;;) {
        const uint32_t t = next_item(q, &item_slot);
        if (t >= total) break;
        const int v = view_of(tb, nv, t);
        const int lt = (int)(t - tb[v]);
        backward_tile<PPT, BB>(tile_view(a, v), d, lt * nv + v, lt);
    }
}

// Kernel variants by (pixels per thread, LDS batch depth), picked at launch
template <int FB>
static void forward_one(const RasterArgs &a, int nt, hipStream_t s) {
    raster_forward_kernel<FB><<<nt, 256, 0, s>>>(a);
}

template <int FB>
static void forward_chunked(const RasterArgs &a, int nt, hipStream_t s) {
    const int g = nt * a.chunks;
    const dim3 pix((a.width + 15) / 16, (a.height + 15) / 16);
    raster_forward_chunk_kernel<FB, true><<<g, 256, 0, s>>>(a);
    raster_chunk_prefix_kernel<<<pix, 256, 0, s>>>(a);
    raster_forward_chunk_kernel<FB, false><<<g, 256, 0, s>>>(a);
    raster_chunk_merge_kernel<<<pix, 256, 0, s>>>(a);
}

void launch_raster_forward(const RasterArgs &args, const Tuning &t, hipStream_t s) {
    RasterArgs a = args;
    a.cull = t.cull;
    const int nt = a.tiles_x * a.tiles_y;
    if (nt <= 0) return;
    if (a.aux) {  // one variant: the prior-guided path is not the tuned one
        raster_forward_kernel<256, true><<<nt, 256, 0, s>>>(a);
        return;
    }
    if (a.chunks > 1) {
        if (t.fwd_batch == 64) forward_chunked<64>(a, nt, s);
        else if (t.fwd_batch == 128) forward_chunked<128>(a, nt, s);
        else forward_chunked<256>(a, nt, s);
        return;
    }
    if (t.fwd_batch == 64) forward_one<64>(a, nt, s);
    else if (t.fwd_batch == 128) forward_one<128>(a, nt, s);
    else forward_one<256>(a, nt, s);
}

template <int PPT>
static void backward_ppt(const RasterArgs &a, int nt, int bb, hipStream_t s) {
    if (a.aux) {
        raster_backward_kernel<PPT, 64, true><<<nt, 256 / PPT, 0, s>>>(a);
        return;
    }
    if (a.chunks > 1) {
        const int g = nt * a.chunks;
        raster_chunk_suffix_kernel<<<dim3((a.width + 15) / 16, (a.height + 15) / 16), 256, 0, s>>>(a);
        if (bb == 32) raster_backward_chunk_kernel<PPT, 32><<<g, 256 / PPT, 0, s>>>(a);
        else if (bb == 64) raster_backward_chunk_kernel<PPT, 64><<<g, 256 / PPT, 0, s>>>(a);
        else raster_backward_chunk_kernel<PPT, 128><<<g, 256 / PPT, 0, s>>>(a);
        return;
    }
    if (bb == 32) raster_backward_kernel<PPT, 32><<<nt, 256 / PPT, 0, s>>>(a);
    else if (bb == 64) raster_backward_kernel<PPT, 64><<<nt, 256 / PPT, 0, s>>>(a);
    else raster_backward_kernel<PPT, 128><<<nt, 256 / PPT, 0, s>>>(a);
}

void launch_raster_backward(const RasterArgs &args, const Tuning &t, hipStream_t s) {
    RasterArgs a = args;
    a.cull = t.cull;
    const int nt = a.tiles_x * a.tiles_y;
    if (nt <= 0) return;
    if (t.bwd_ppt == 4) backward_ppt<4>(a, nt, t.bwd_batch, s);
    else if (t.bwd_ppt == 2) backward_ppt<2>(a, nt, t.bwd_batch, s);
    else backward_ppt<1>(a, nt, t.bwd_batch, s);
}

// Depth chunks per tile for a one-view frame of nt tiles and m intersections: with
// fewer than 400 tiles (the reference's own limit for its chunked path), chunks of at
// least 16 intersections on average, up to about 16384 workgroups, between 2 and 64;
// else 1. Measured on 128x128 (64 tiles) and 320x240 (300 tiles) views with 280 to
// 7000 intersections per tile: the step time falls with the chunk count until the
// chunks hold about 16 intersections or the chunk planes' traffic takes over.
int auto_chunks(int nt, int64_t m) {
    if (nt <= 0 || nt >= 400 || m < 32 * (int64_t)nt) return 1;
    const int64_t by_length = m / (16 * (int64_t)nt), by_groups = (16384 + nt - 1) / nt;
    return (int)std::max<int64_t>(2, std::min<int64_t>(std::min<int64_t>(by_length, by_groups), 64));
}

namespace {

// Function-local statics: initialised once, thread-safely, on first use (the render
// server launches frames from several host threads).
int device_cus() {
    static const int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        return n;
    }();
    return cus;
}

template <typename K>
int resident_blocks(K kernel, int threads) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    return per_cu * device_cus();
}

}  // namespace

template <int FB>
static int fwd_grid() {
    static const int g = resident_blocks(raster_forward_batch_kernel<FB>, 256);
    return g;
}

template <int PPT, int BB>
static int bwd_grid() {
    static const int g = resident_blocks(raster_backward_batch_kernel<PPT, BB>, 256 / PPT);
    return g;
}

template <int PPT>
static int bwd_grid_ppt(int bb) {
    return bb == 32 ? bwd_grid<PPT, 32>() : bb == 64 ? bwd_grid<PPT, 64>() : bwd_grid<PPT, 128>();
}

int batch_raster_grid(bool backward, const Tuning &t) {
    if (!backward) return t.fwd_batch == 64 ? fwd_grid<64>() : t.fwd_batch == 128 ? fwd_grid<128>() : fwd_grid<256>();
    return t.bwd_ppt == 4 ? bwd_grid_ppt<4>(t.bwd_batch) : t.bwd_ppt == 2 ? bwd_grid_ppt<2>(t.bwd_batch)
                                                                          : bwd_grid_ppt<1>(t.bwd_batch);
}

__global__ void zero_heads(uint32_t *p) {
    if (threadIdx.x < NUM_XCD) p[threadIdx.x] = 0u;
}

void launch_batch_raster_forward(const BatchRasterArgs &args, const Tuning &t, hipStream_t s) {
    BatchRasterArgs a = args;
    a.cull = t.cull;
    zero_heads<<<1, 64, 0, s>>>(a.queue);
    const int g = batch_raster_grid(false, t);
    if (t.fwd_batch == 64) raster_forward_batch_kernel<64><<<g, 256, 0, s>>>(a);
    else if (t.fwd_batch == 128) raster_forward_batch_kernel<128><<<g, 256, 0, s>>>(a);
    else raster_forward_batch_kernel<256><<<g, 256, 0, s>>>(a);
}

template <int PPT>
static void batch_backward_ppt(const BatchRasterArgs &a, int g, int bb, hipStream_t s) {
    if (bb == 32) raster_backward_batch_kernel<PPT, 32><<<g, 256 / PPT, 0, s>>>(a);
    else if (bb == 64) raster_backward_batch_kernel<PPT, 64><<<g, 256 / PPT, 0, s>>>(a);
    else raster_backward_batch_kernel<PPT, 128><<<g, 256 / PPT, 0, s>>>(a);
}

void launch_batch_raster_backward(const BatchRasterArgs &args, const Tuning &t, hipStream_t s) {
    BatchRasterArgs a = args;
    a.cull = t.cull;
    zero_heads<<<1, 64, 0, s>>>(a.queue + NUM_XCD);
    const int g = batch_raster_grid(true, t);
    if (t.bwd_ppt == 4) batch_backward_ppt<4>(a, g, t.bwd_batch, s);
    else if (t.bwd_ppt == 2) batch_backward_ppt<2>(a, g, t.bwd_batch, s);
    else batch_backward_ppt<1>(a, g, t.bwd_batch, s);
}

}  // namespace msplat
