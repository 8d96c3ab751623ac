// Rasterizer kernels: one workgroup per tile for a single view, and persistent
// workgroups that pull (view, tile) items from a device work queue for a batch of views
// of mixed resolution. Both run the tile bodies of raster_tile.h.
#include "raster_tile.h"

namespace msplat {

namespace {

__device__ inline TileData tile_data(const RasterArgs &a) {
    return TileData{a.tile_start, a.gid, a.splat_a, a.splat_b, a.splat_blue, a.rect, a.offsets,
                    a.slot_a,     a.slot_b, a.slot_c};
}

__device__ inline TileView tile_view(const RasterArgs &a) {
    return TileView{a.width, a.height, a.tiles_x, {a.bg[0], a.bg[1], a.bg[2]}, a.image, a.final_T, a.final_idx,
                    a.v_image};
}

__device__ inline TileData tile_data(const BatchRasterArgs &a) {
    return TileData{a.tile_start, a.gid, a.splat_a, a.splat_b, a.splat_blue, a.rect, a.offsets,
                    a.slot_a,     a.slot_b, a.slot_c};
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

// Next item of the queue for the whole workgroup (one atomic per item)
__device__ inline uint32_t next_item(uint32_t *queue, uint32_t *slot) {
    if (threadIdx.x == 0) *slot = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t it = *slot;
    __syncthreads();
    return it;
}

}  // namespace

__global__ __launch_bounds__(256) void raster_forward_kernel(RasterArgs a) {
    forward_tile(tile_view(a), tile_data(a), blockIdx.x, blockIdx.x);
}

template <int PPT>
__global__ __launch_bounds__(256 / PPT) void raster_backward_kernel(RasterArgs a) {
    backward_tile<PPT>(tile_view(a), tile_data(a), blockIdx.x, blockIdx.x);
}

// Persistent forward over every (view, tile) of a batch. Every workgroup loops until
// the queue passes the batch's tile count, so the grid drains on its own.
__global__ __launch_bounds__(256) void raster_forward_batch_kernel(BatchRasterArgs a) {
    __shared__ uint32_t tb[MAX_VIEWS + 1];
    __shared__ uint32_t item_slot;
    const int nv = a.num_views;
    for (int i = threadIdx.x; i <= nv; i += 256) tb[i] = a.tab->tile_base[i];
    const uint32_t total = a.tab->total_tiles;
    const TileData d = tile_data(a);
    for (;;) {
        const uint32_t t = next_item(a.queue, &item_slot);
        if (t >= total) break;
        const int v = view_of(tb, nv, t);
        forward_tile(tile_view(a, v), d, (int)t, (int)(t - tb[v]));
    }
}

template <int PPT>
__global__ __launch_bounds__(256 / PPT) void raster_backward_batch_kernel(BatchRasterArgs a) {
    __shared__ uint32_t tb[MAX_VIEWS + 1];
    __shared__ uint32_t item_slot;
    const int nv = a.num_views;
    for (int i = threadIdx.x; i <= nv; i += 256 / PPT) tb[i] = a.tab->tile_base[i];
    const uint32_t total = a.tab->total_tiles;
    const TileData d = tile_data(a);
    for (;;) {
        const uint32_t t = next_item(a.queue + 1, &item_slot);
        if (t >= total) break;
        const int v = view_of(tb, nv, t);
        backward_tile<PPT>(tile_view(a, v), d, (int)t, (int)(t - tb[v]));
    }
}

void launch_raster_forward(const RasterArgs &a, hipStream_t s) {
    const int nt = a.tiles_x * a.tiles_y;
    if (nt > 0) raster_forward_kernel<<<nt, 256, 0, s>>>(a);
}

void launch_raster_backward(const RasterArgs &a, int ppt, hipStream_t s) {
    const int nt = a.tiles_x * a.tiles_y;
    if (nt <= 0) return;
    if (ppt == 4)
        raster_backward_kernel<4><<<nt, 64, 0, s>>>(a);
    else if (ppt == 2)
        raster_backward_kernel<2><<<nt, 128, 0, s>>>(a);
    else
        raster_backward_kernel<1><<<nt, 256, 0, s>>>(a);
}

namespace {

int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
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

int batch_raster_grid(bool backward, int ppt) {
    static int g[4] = {0, 0, 0, 0};
    const int i = !backward ? 0 : ppt == 4 ? 3 : ppt == 2 ? 2 : 1;
    if (!g[i]) {
        if (i == 0) g[i] = resident_blocks(raster_forward_batch_kernel, 256);
        else if (i == 1) g[i] = resident_blocks(raster_backward_batch_kernel<1>, 256);
        else if (i == 2) g[i] = resident_blocks(raster_backward_batch_kernel<2>, 128);
        else g[i] = resident_blocks(raster_backward_batch_kernel<4>, 64);
    }
    return g[i];
}

void launch_batch_raster_forward(const BatchRasterArgs &a, hipStream_t s) {
    fill_u32<<<1, 1, 0, s>>>(a.queue, 0u);
    raster_forward_batch_kernel<<<batch_raster_grid(false, 1), 256, 0, s>>>(a);
}

void launch_batch_raster_backward(const BatchRasterArgs &a, int ppt, hipStream_t s) {
    fill_u32<<<1, 1, 0, s>>>(a.queue + 1, 0u);
    if (ppt == 4)
        raster_backward_batch_kernel<4><<<batch_raster_grid(true, 4), 64, 0, s>>>(a);
    else if (ppt == 2)
        raster_backward_batch_kernel<2><<<batch_raster_grid(true, 2), 128, 0, s>>>(a);
    else
        raster_backward_batch_kernel<1><<<batch_raster_grid(true, 1), 256, 0, s>>>(a);
}

}  // namespace msplat
