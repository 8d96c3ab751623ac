// Projection + SH colour, binning of the splats into 16x16 tiles, stable LSD radix
// sort and per-tile range extraction.
//
// Layout in HBM (per record, SoA, written once per view; a record is a gaussian in one
// view, record v * n + g in a batch of views):
//   splat_a = (x, y, conic.x, conic.y)   float4
//   splat_b = (conic.z, sigmoid(opacity), r, g)   float4
//   splat_blue = b, depths, aabb (float2), rect (x0 | x1 << 16, y0 | y1 << 16), count
//
// Binning keeps the reference's per-tile order, ascending (depth bits, gaussian id)
// (core/cpu/msplat_cpu.cpp:505-537), with two small stable sorts in place of one sort
// of 64-bit (tile | depth) keys:
//   1. the visible records are compacted in id order and stably sorted by their 32-bit
//      depth bits (4 passes over the visible records, not over the intersections);
//   2. each record's intersections are emitted in that depth order with its tile id as
//      a 32-bit key, and a stable sort by tile (ceil(tile bits / 8) passes) leaves every
//      tile's intersections in (depth, id) order.
// Records own contiguous slot ranges (offsets[r], count[r]) in emission order, tiles
// row-major within a record, which the backward's per-record reductions read.
#include <algorithm>

#include "kernels.h"
#include "proj_math.h"

namespace msplat {

// Block-wide exclusive scan of one u32 per thread (256 threads, 4 waves).
__device__ inline uint32_t block_scan_u32(uint32_t v, uint32_t &total, uint32_t *lds4) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds4[w] = x;
    __syncthreads();
    uint32_t pre = 0;
    total = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (i < w) pre += lds4[i];
        total += lds4[i];
    }
    __syncthreads();
    return pre + x - v;
}

// ── projection + SH colour (project_and_sh, msplat_cpu.cpp:452-500) ────────────
// Also writes the workgroup's intersection count and visible-record count.
// SP: the camera comes from the device StepParams block (a captured step), chosen at
// compile time so that the launch-by-launch step reads it from the kernel arguments.
// The SH coefficients of a wave's visible gaussians are staged in LDS by stage_sh
// (dynamic LDS: project_sh_lds bytes for the degree in use).
template <bool SP>
__global__ __launch_bounds__(256) void project_kernel(ProjectArgs a) {
    extern __shared__ float shl[];
    __shared__ uint32_t lds4[4];
    __shared__ uint8_t widx[4][64];
    const msplat_hip_camera &cam = SP ? a.sp->cam : a.cam;
    const int g = blockIdx.x * 256 + threadIdx.x;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t count = 0;
    bool need = false;
    float p[3] = {0.f, 0.f, 0.f}, pv[3] = {0.f, 0.f, 0.f}, conic[3] = {0.f, 0.f, 0.f}, radius = 0.f;
    float px = 0.f, py = 0.f, ax = 0.f, ay = 0.f;
    int x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    if (g < a.n) {
        a.radii[g] = 0;
        a.count[g] = 0;
        p[0] = a.means[3 * g]; p[1] = a.means[3 * g + 1]; p[2] = a.means[3 * g + 2];
        cam_space(cam.viewmat, p, pv);
        if (pv[2] > cam.clip_thresh) {
            const float s[3] = {exp_cr(a.scales[3 * g]), exp_cr(a.scales[3 * g + 1]), exp_cr(a.scales[3 * g + 2])};
            const float q[4] = {a.quats[4 * g], a.quats[4 * g + 1], a.quats[4 * g + 2], a.quats[4 * g + 3]};
            float cov[6], c2[3];
            covariance3d(s, cam.glob_scale, q, cov);
            const float tfx = 0.5f * (float)cam.width / cam.fx, tfy = 0.5f * (float)cam.height / cam.fy;
            ewa_cov2d(cov, cam.viewmat, cam.fx, cam.fy, tfx, tfy, pv, c2);
            if (conic_radius(c2, conic, radius)) {
                to_pixel(cam.projmat, p, cam.width, cam.height, cam.cx, cam.cy, px, py);
                ax = ceilf(3.f * sqrtf(c2[0]));
                ay = ceilf(3.f * sqrtf(c2[2]));
                tile_rect(px, py, ax, ay, cam.tiles_x, cam.tiles_y, x0, y0, x1, y1);
                need = (x1 - x0) * (y1 - y0) > 0;
            }
        }
    }
    // staged when higher bands are in use, up to SH degree 3 (project_sh_lds); the DC
    // alone (three floats a lane) and degree 4 are read in place
    const size_t stride = 3 * (size_t)(sh_bases(cam.degree) - 1);
    const bool staged = cam.degree <= 3 && cam.degrees_to_use >= 1;
    const int u = 3 * (int)sh_bases(cam.degrees_to_use), us = u | 1;
    float *wl = shl + (size_t)w * 64 * us;
    if (staged) stage_sh(need, g - lane, a.dc, a.rest, stride, u, us, wl, widx[w]);
    if (need) {
        float d[3], col[3];
        view_direction(p, cam.cam_pos, d);
        const float *mine = wl + lane * us;
        if (staged) sh_color(cam.degrees_to_use, d, mine, mine + 3, col);
        else sh_color(cam.degrees_to_use, d, a.dc + 3 * (size_t)g, a.rest + stride * g, col);
        const float opac = 1.f / (1.f + exp_cr(-a.opac[g]));
        a.splat_a[g] = make_float4(px, py, conic[0], conic[1]);
        a.splat_b[g] = make_float4(conic[2], opac, col[0], col[1]);
        a.splat_blue[g] = col[2];
        a.depths[g] = pv[2];
        a.aabb[g] = make_float2(ax, ay);
        a.rect[g] = make_uint2((uint32_t)x0 | ((uint32_t)x1 << 16), (uint32_t)y0 | ((uint32_t)y1 << 16));
        a.radii[g] = (int)radius;
        count = (uint32_t)((x1 - x0) * (y1 - y0));
        a.count[g] = count;
    }
    uint32_t total, vis;
    block_scan_u32(count, total, lds4);
    block_scan_u32(count > 0 ? 1u : 0u, vis, lds4);
    if (threadIdx.x == 0) {
        a.block_sums[blockIdx.x] = total;
        a.block_vis[blockIdx.x] = vis;
    }
}

// Tile rectangles and counts re-derived from a frame's projection outputs (pixel
// centre in splat_a, aabb, radii), for binning outputs a caller injected.
__global__ __launch_bounds__(256) void rect_count_kernel(int n, const float4 *splat_a, const float2 *aabb,
                                                        const int32_t *radii, int tx, int ty, uint2 *rect,
                                                        uint32_t *count, uint32_t *block_sums, uint32_t *block_vis) {
    __shared__ uint32_t lds4[4];
    const int g = blockIdx.x * 256 + threadIdx.x;
    uint32_t c = 0;
    if (g < n) {
        if (radii[g] > 0) {
            const float4 A = splat_a[g];
            const float2 e = aabb[g];
            int x0, y0, x1, y1;
            tile_rect(A.x, A.y, e.x, e.y, tx, ty, x0, y0, x1, y1);
            if ((x1 - x0) * (y1 - y0) > 0) c = (uint32_t)((x1 - x0) * (y1 - y0));
            rect[g] = make_uint2((uint32_t)x0 | ((uint32_t)x1 << 16), (uint32_t)y0 | ((uint32_t)y1 << 16));
        }
        count[g] = c;
    }
    uint32_t total, vis;
    block_scan_u32(c, total, lds4);
    block_scan_u32(c > 0 ? 1u : 0u, vis, lds4);
    if (threadIdx.x == 0) {
        block_sums[blockIdx.x] = total;
        block_vis[blockIdx.x] = vis;
    }
}

// Exclusive scan of the per-block counts of a small array (one block, sequential
// chunks); writes the bases and the total (u64, so that > 2^31 is detected).
__global__ __launch_bounds__(256) void scan_blocks_kernel(uint32_t *sums, int nblocks, uint64_t *total,
                                                          uint32_t *total32, uint32_t *zero_also) {
    __shared__ uint32_t lds4[4];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) {
        carry = 0;
        if (zero_also) *zero_also = 0;
    }
    __syncthreads();
    for (int base = 0; base < nblocks; base += 256) {
        const int i = base + threadIdx.x;
        uint32_t v = i < nblocks ? sums[i] : 0, t;
        uint32_t ex = block_scan_u32(v, t, lds4);
        uint64_t c = carry;
        if (i < nblocks) sums[i] = (uint32_t)(c + ex);
        __syncthreads();
        if (threadIdx.x == 0) carry = c + t;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *total = carry;
        *total32 = carry > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)carry;
    }
}

// ── multi-level exclusive scan of a u32 array (4096 entries per workgroup) ─────
constexpr int SCAN_CHUNK = 4096;

__global__ __launch_bounds__(256) void scan_partial_kernel(const uint32_t *a, uint32_t L, uint32_t *partial) {
    __shared__ uint32_t lds4[4];
    const uint32_t base = blockIdx.x * SCAN_CHUNK + threadIdx.x * 16;
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
        if (base + i < L) s += a[base + i];
    uint32_t t;
    block_scan_u32(s, t, lds4);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Scans the chunk totals; the grand total goes to total/total32. Over `cap`, total32 is
// 0 and *overflow is set (when overflow is given; with `applied`, a set overflow also
// counts as over capacity).
__global__ __launch_bounds__(256) void scan_top_kernel(uint32_t *partial, int nchunks, ScanOut o) {
    __shared__ uint32_t lds4[4];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nchunks; base += 256) {
        const int i = base + threadIdx.x;
        uint32_t v = i < nchunks ? partial[i] : 0, t;
        const uint32_t ex = block_scan_u32(v, t, lds4);
        const uint64_t c = carry;
        if (i < nchunks) partial[i] = (uint32_t)(c + ex);
        __syncthreads();
        if (threadIdx.x == 0) carry = c + t;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint64_t tot = carry;
        const bool over = tot > (uint64_t)o.cap || (o.applied && *o.overflow);
        if (o.total) *o.total = tot;
        if (o.total32) *o.total32 = over ? (o.overflow ? 0u : 0xFFFFFFFFu) : (uint32_t)tot;
        if (o.overflow) *o.overflow = over ? 1u : 0u;
        if (o.zero_also) *o.zero_also = 0u;
        if (o.applied && !over) *o.applied += 1u;
        if (o.running) *o.running += tot;
    }
}

// The whole scan in one workgroup, 4096 entries at a time (small arrays)
__device__ inline void scan_small(uint32_t *a, uint32_t L, const ScanOut &o) {
    __shared__ uint32_t lds4[4];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t chunk = 0; chunk < L; chunk += SCAN_CHUNK) {
        const uint32_t base = chunk + threadIdx.x * 16;
        uint32_t v[16], sum = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            v[i] = base + i < L ? a[base + i] : 0u;
            sum += v[i];
        }
        uint32_t t;
        const uint32_t ex = block_scan_u32(sum, t, lds4);
        const uint64_t c = carry;
        uint32_t run = (uint32_t)c + ex;
#pragma unroll
        for (int i = 0; i < 16; i++)
            if (base + i < L) {
                a[base + i] = run;
                run += v[i];
            }
        __syncthreads();
        if (threadIdx.x == 0) carry = c + t;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint64_t tot = carry;
        const bool over = tot > (uint64_t)o.cap || (o.applied && *o.overflow);
        if (o.total) *o.total = tot;
        if (o.total32) *o.total32 = over ? (o.overflow ? 0u : 0xFFFFFFFFu) : (uint32_t)tot;
        if (o.overflow) *o.overflow = over ? 1u : 0u;
        if (o.zero_also) *o.zero_also = 0u;
        if (o.applied && !over) *o.applied += 1u;
        if (o.running) *o.running += tot;
    }
}

__global__ __launch_bounds__(256) void scan_small_kernel(uint32_t *a, uint32_t L, ScanOut o) { scan_small(a, L, o); }

// Two small scans of the same length in one launch, one workgroup each
__global__ __launch_bounds__(256) void scan_small_pair_kernel(uint32_t *a0, ScanOut o0, uint32_t *a1, ScanOut o1,
                                                              uint32_t L) {
    if (blockIdx.x == 0) scan_small(a0, L, o0);
    else scan_small(a1, L, o1);
}

__global__ __launch_bounds__(256) void scan_apply_kernel(uint32_t *a, uint32_t L, const uint32_t *partial) {
    __shared__ uint32_t lds4[4];
    const uint32_t base = blockIdx.x * SCAN_CHUNK + threadIdx.x * 16;
    uint32_t v[16], s = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        v[i] = base + i < L ? a[base + i] : 0u;
        s += v[i];
    }
    uint32_t t;
    uint32_t ex = block_scan_u32(s, t, lds4) + partial[blockIdx.x];
#pragma unroll
    for (int i = 0; i < 16; i++)
        if (base + i < L) {
            a[base + i] = ex;
            ex += v[i];
        }
}

// ── visible records in id order, keyed by depth bits ──────────────────────────
__global__ __launch_bounds__(256) void compact_kernel(CompactArgs a) {
    __shared__ uint32_t lds4[4];
    const int v = blockIdx.y;
    const int g = blockIdx.x * 256 + threadIdx.x;
    const size_t r = (size_t)v * a.n + g;
    const bool vis = g < a.n && a.count[r] > 0;
    uint32_t t;
    const uint32_t ex = block_scan_u32(vis ? 1u : 0u, t, lds4);
    if (vis) {
        const uint32_t pos = a.block_vis[(size_t)v * a.nbx + blockIdx.x] + ex;
        a.dkeys[pos] = __float_as_uint(a.depths[r]);
        a.dvals[pos] = (int32_t)r;
    }
}

// ── stable LSD radix sort, 8-bit digits, 32- or 64-bit keys, int32 payload ─────
// Each workgroup owns an element block of ITEMS consecutive elements (4096 for large
// sorts, 1024 for small ones, so that they still spread over the device): an upsweep
// histogram, a per-digit scan over blocks, and a scatter that ranks the elements stably
// within the block (wave multi-split over ballots of the digit bits), stages them in
// LDS in digit order, and writes each digit's run to consecutive addresses.

// The workgroups of a pass loop over the element blocks that the device-side count
// needs, so a grid sized for a capacity does no empty work; histogram column b of
// digit d lives at hist[d * stride + b].
template <int ITEMS>
__device__ inline uint32_t rs_blocks(const uint32_t *n_ptr) { return (*n_ptr + ITEMS - 1) / ITEMS; }

template <typename K, int ITEMS>
__global__ __launch_bounds__(256) void rs_hist_kernel(const K *keys, const uint32_t *n_ptr, int shift, uint32_t mask,
                                                     uint32_t *hist, int stride) {
    __shared__ uint32_t h[256];
    const uint32_t n = *n_ptr, nb = rs_blocks<ITEMS>(n_ptr);
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        h[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t start = b * ITEMS;
        for (int r = 0; r < ITEMS / 256; r++) {
            uint32_t e = start + r * 256 + threadIdx.x;
            if (e < n) atomicAdd(&h[(uint32_t)(keys[e] >> shift) & mask], 1u);
        }
        __syncthreads();
        hist[(size_t)threadIdx.x * stride + b] = h[threadIdx.x];
        __syncthreads();
    }
}

// one block per digit: exclusive scan of that digit's counts over blocks, digit total;
// each thread takes SCAN_RUN consecutive counts, so 2048 blocks take one round
constexpr int SCAN_RUN = 8;
template <int ITEMS>
__global__ __launch_bounds__(256) void rs_scan_kernel(uint32_t *hist, int stride, const uint32_t *n_ptr,
                                                     uint32_t *digit_total) {
    __shared__ uint32_t lds4[4];
    __shared__ uint32_t carry;
    const int nblocks = (int)rs_blocks<ITEMS>(n_ptr);
    uint32_t *row = hist + (size_t)blockIdx.x * stride;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nblocks; base += 256 * SCAN_RUN) {
        const int i0 = base + threadIdx.x * SCAN_RUN;
        uint32_t v[SCAN_RUN], sum = 0, t;
#pragma unroll
        for (int j = 0; j < SCAN_RUN; j++) {
            v[j] = i0 + j < nblocks ? row[i0 + j] : 0u;
            sum += v[j];
        }
        const uint32_t ex = block_scan_u32(sum, t, lds4);
        const uint32_t c = carry;
        uint32_t run = c + ex;
#pragma unroll
        for (int j = 0; j < SCAN_RUN; j++)
            if (i0 + j < nblocks) {
                row[i0 + j] = run;
                run += v[j];
            }
        __syncthreads();
        if (threadIdx.x == 0) carry = c + t;
        __syncthreads();
    }
    if (threadIdx.x == 0) digit_total[blockIdx.x] = carry;
}

// Wave w ranks its own contiguous quarter of the workgroup's elements, 64 at a time,
// with private per-digit counters (no workgroup barrier inside the loop); the waves'
// counts are then combined in wave order, which keeps the sort stable.
// One element block of a pass; `my_base` is this thread's digit's global start for the block.
template <typename K, int ITEMS>
__device__ inline void rs_scatter_block(const K *kin, const int32_t *vin, K *kout, int32_t *vout, uint32_t n,
                                        int shift, uint32_t mask, uint32_t start, uint32_t my_base) {
    constexpr int R = ITEMS / 256;  // rounds of 64 elements per wave
    __shared__ K sk[ITEMS];
    __shared__ int32_t sv[ITEMS];
    __shared__ uint32_t gbase[256], lstart[256];
    __shared__ uint32_t wcnt[4][256];
    __shared__ uint32_t lds4[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    gbase[tid] = my_base;
#pragma unroll
    for (int i = 0; i < 4; i++) wcnt[w][lane + 64 * i] = 0;
    const uint32_t wstart = start + (uint32_t)w * (64 * R);
    K key[R];
    int32_t val[R];
    const bool has_vals = vin != nullptr;  // null: the keys carry their payload
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t e = wstart + r * 64 + lane;
        key[r] = e < n ? kin[e] : K(0);
        val[r] = has_vals && e < n ? vin[e] : 0;
    }
    wave_sync();
    // per element: digit (bits 24..31) and rank among the wave's elements of that digit
    uint32_t dr[R];
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const bool valid = wstart + r * 64 + lane < n;
        const uint32_t digit = valid ? (uint32_t)(key[r] >> shift) & mask : 0;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (digit >> b) & 1;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t before = valid ? wcnt[w][digit] : 0u;
        wave_sync();
        if (valid && (peers & lt) == 0) wcnt[w][digit] = before + (uint32_t)__popcll(peers);
        wave_sync();
        dr[r] = (digit << 24) | (before + (uint32_t)__popcll(peers & lt));
    }
    __syncthreads();
    {   // per digit: wave offsets (in wave order) and the block's local start
        const uint32_t c0 = wcnt[0][tid], c1 = wcnt[1][tid], c2 = wcnt[2][tid], c3 = wcnt[3][tid];
        uint32_t t;
        const uint32_t ls = block_scan_u32(c0 + c1 + c2 + c3, t, lds4);
        lstart[tid] = ls;
        wcnt[0][tid] = ls;
        wcnt[1][tid] = ls + c0;
        wcnt[2][tid] = ls + c0 + c1;
        wcnt[3][tid] = ls + c0 + c1 + c2;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (wstart + r * 64 + lane < n) {
            const uint32_t p = wcnt[w][dr[r] >> 24] + (dr[r] & 0xFFFFFF);
            sk[p] = key[r];
            if (has_vals) sv[p] = val[r];
        }
    }
    __syncthreads();
    const uint32_t cnt = min((uint32_t)ITEMS, n - start);
    for (uint32_t p = tid; p < cnt; p += 256) {
        const K k = sk[p];
        const uint32_t d = (uint32_t)(k >> shift) & mask;
        const uint32_t o = gbase[d] + (p - lstart[d]);
        kout[o] = k;
        if (has_vals) vout[o] = sv[p];
    }
}


template <typename K, int ITEMS>
__global__ __launch_bounds__(256) void rs_scatter_kernel(const K *kin, const int32_t *vin, K *kout, int32_t *vout,
                                                        const uint32_t *n_ptr, int shift, uint32_t mask,
                                                        const uint32_t *hist, int stride, const uint32_t *digit_total) {
    __shared__ uint32_t dex[256];
    __shared__ uint32_t lds4[4];
    {   // exclusive prefix of the digit totals
        uint32_t t;
        dex[threadIdx.x] = block_scan_u32(digit_total[threadIdx.x], t, lds4);
    }
    const uint32_t nb = rs_blocks<ITEMS>(n_ptr);
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        __syncthreads();
        rs_scatter_block<K, ITEMS>(kin, vin, kout, vout, *n_ptr, shift, mask, b * ITEMS,
                                   dex[threadIdx.x] + hist[(size_t)threadIdx.x * stride + b]);
    }
}

// ── intersections in depth order ──────────────────────────────────────────────

// Per 256-entry block of the depth-ordered list: the sum of its records' counts
// (0 for blocks past the list's device-side length). Each record's count also goes to
// lcount[i] in list order, so the offsets kernel reads it without a second gather.
__global__ __launch_bounds__(256) void order_sums_kernel(const int32_t *list, const uint32_t *nvis_ptr,
                                                        const uint32_t *count, uint32_t *sums, uint32_t nchunks_max,
                                                        uint32_t *lcount) {
    __shared__ uint32_t lds4[4];
    const uint32_t nvis = *nvis_ptr;
    for (uint32_t chunk = blockIdx.x; chunk < nchunks_max; chunk += gridDim.x) {
        if (chunk * 256 >= nvis) {  // past the list: zero sums, nothing to read
            if (threadIdx.x == 0) sums[chunk] = 0;
            continue;
        }
        const uint32_t i = chunk * 256 + threadIdx.x;
        const uint32_t c = i < nvis ? count[list[i]] : 0u;
        if (i < nvis) lcount[i] = c;
        uint32_t t;
        block_scan_u32(c, t, lds4);
        if (threadIdx.x == 0) sums[chunk] = t;
    }
}

// Record list[i] writes its tiles in row-major order at offsets[r] = the exclusive
// prefix of the counts in list order (also kept in list order, loff[i], for the
// emission), and joins the big list when it has more than BIG_SLOTS intersections.
// loff[i] holds the record's count on entry (order_sums_kernel).
__global__ __launch_bounds__(256) void keygen_offsets_kernel(KeygenArgs a) {
    __shared__ uint32_t lds4[4];
    __shared__ uint32_t s_bigbase;
    if (a.overflow && *a.overflow) return;
    const uint32_t nvis = *a.nvis_ptr;
    for (uint32_t chunk = blockIdx.x; chunk * 256 < nvis; chunk += gridDim.x) {
        const uint32_t i = chunk * 256 + threadIdx.x;
        const int32_t r = i < nvis ? a.list[i] : -1;
        const uint32_t c = r >= 0 ? a.loff[i] : 0u;
        uint32_t total;
        uint32_t base;
        if (a.chunk_sums) {  // the sums of the chunks before this one, added here
            uint32_t part = 0;
            for (uint32_t q = threadIdx.x; q < chunk; q += 256) part += a.chunk_sums[q];
            block_scan_u32(part, base, lds4);
        } else {
            base = a.block_bases[chunk];
        }
        const uint32_t off = base + block_scan_u32(c, total, lds4);
        if (r >= 0) {
            a.offsets[r] = off;
            a.loff[i] = off;
        }
        uint32_t nbig;
        const uint32_t bex = block_scan_u32(c > BIG_SLOTS ? 1u : 0u, nbig, lds4);
        if (threadIdx.x == 0 && nbig) s_bigbase = atomicAdd(a.big_count, nbig);
        __syncthreads();
        if (c > BIG_SLOTS) a.big_list[s_bigbase + bex] = (uint32_t)r;
        __syncthreads();
    }
}

// Key emission balanced over the output: workgroup b writes intersections
// [b * EMIT_CHUNK, (b + 1) * EMIT_CHUNK). It finds the records covering its slice by a
// block search over loff and stages their offsets, ids and tile rectangles in LDS, marking
// where each record after the first starts. Thread t then writes the slice's outputs
// 8t..8t+7: the record of output p is the number of marks at or before p (a block scan of
// the threads' mark counts), and its tile follows from the previous output's by one step
// along the rectangle's row, so one division per thread places its first output. The eight
// keys (and ids) go out as two 16-byte stores; consecutive threads write consecutive
// keys whatever the records' footprints (depth order puts the large near splats next to
// each other).
constexpr int EMIT_ITEMS = 8;
constexpr int EMIT_CHUNK = 256 * EMIT_ITEMS;

template <bool BATCH>
__global__ __launch_bounds__(256) void keygen_emit_kernel(KeygenArgs a) {
    __shared__ uint32_t s_off[EMIT_CHUNK];
    __shared__ int32_t s_rec[EMIT_CHUNK];
    __shared__ uint2 s_rect[EMIT_CHUNK];
    __shared__ __align__(8) uint8_t s_mark[EMIT_CHUNK];
    __shared__ uint32_t s_first, s_n;
    __shared__ uint32_t lds4[4];
    if (a.overflow && *a.overflow) return;
    const uint32_t M = *a.m_ptr;
    const uint32_t o0 = blockIdx.x * EMIT_CHUNK;
    if (o0 >= M) return;
    const uint32_t o1 = min(o0 + (uint32_t)EMIT_CHUNK, M);
    {
        // records first..last cover [o0, o1): the last i with loff[i] <= o0, o1 - 1 (loff
        // rises from loff[0] = 0). A 256-ary search: each round every thread tests one
        // probe of each range, and the count of probes at or below the target narrows
        // the range [lo, hi) 256-fold (3 rounds for 3M records, every thread in step).
        const uint32_t nvis = *a.nvis_ptr;
        uint32_t lo1 = 0, hi1 = nvis, lo2 = 0, hi2 = nvis;
        while (hi1 - lo1 > 1 || hi2 - lo2 > 1) {
            const uint32_t st1 = (hi1 - lo1 + 255) / 256, st2 = (hi2 - lo2 + 255) / 256;
            const uint32_t i1 = lo1 + threadIdx.x * st1, i2 = lo2 + threadIdx.x * st2;
            const int c1 = __syncthreads_count(i1 < hi1 && a.loff[i1] <= o0);
            const int c2 = __syncthreads_count(i2 < hi2 && a.loff[i2] <= o1 - 1);
            lo1 += (uint32_t)(c1 - 1) * st1;
            hi1 = min(lo1 + st1, hi1);
            lo2 += (uint32_t)(c2 - 1) * st2;
            hi2 = min(lo2 + st2, hi2);
        }
        if (threadIdx.x == 0) {
            s_first = lo1;
            s_n = lo2 - lo1 + 1;  // <= o1 - o0: every record after the first starts inside the slice
        }
        ((uint32_t *)s_mark)[threadIdx.x] = 0u;
        ((uint32_t *)s_mark)[threadIdx.x + 256] = 0u;
    }
    __syncthreads();
    const uint32_t first = s_first, nrec = s_n;
    for (uint32_t j = threadIdx.x; j < nrec; j += 256) {
        const uint32_t off = a.loff[first + j];
        const int32_t rec = a.list[first + j];
        s_off[j] = off;
        s_rec[j] = rec;
        s_rect[j] = a.rect[rec];
        if (j > 0) s_mark[off - o0] = 1;  // in (0, o1 - o0): records after the first start inside
    }
    __syncthreads();
    const uint32_t p0 = EMIT_ITEMS * threadIdx.x;
    const uint2 mk = *(const uint2 *)(s_mark + p0);
    uint32_t m[EMIT_ITEMS];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        m[i] = (mk.x >> (8 * i)) & 0xFFu;
        m[4 + i] = (mk.y >> (8 * i)) & 0xFFu;
    }
    uint32_t total;
    uint32_t j = block_scan_u32(m[0] + m[1] + m[2] + m[3] + m[4] + m[5] + m[6] + m[7], total, lds4);
    if (o0 + p0 >= o1) return;
    uint32_t kv[EMIT_ITEMS], vv[EMIT_ITEMS];
    uint32_t q = 0, rem = 0, w = 1, x0 = 0, y0 = 0, tiles_x = (uint32_t)a.tiles_x;
    int32_t r = 0;
#pragma unroll
    for (int i = 0; i < EMIT_ITEMS; i++) {
        const bool start = m[i] != 0u;  // record j + 1 starts at this output
        j += m[i];
        if (i == 0 || start) {
            r = s_rec[j];
            const uint2 rr = s_rect[j];
            x0 = rr.x & 0xFFFF;
            w = (rr.x >> 16) - x0;
            y0 = rr.y & 0xFFFF;
            const uint32_t local = o0 + p0 + (uint32_t)i - s_off[j];
            q = local / w;
            rem = local - q * w;
            // local tile id (a batch keeps its views apart by the emission order)
            if (BATCH) tiles_x = (uint32_t)a.tab->views[r / a.n].cam.tiles_x;
        } else if (++rem == w) {
            rem = 0;
            q++;
        }
        const uint32_t tile = (y0 + q) * tiles_x + (x0 + rem);
        if (!BATCH && a.gid_bits) {
            kv[i] = (tile << a.gid_bits) | (uint32_t)r;
        } else {
            kv[i] = tile;
            vv[i] = (uint32_t)r;
        }
    }
    const bool vals = BATCH || !a.gid_bits;
    const uint32_t ob = o0 + p0;
    if (ob + EMIT_ITEMS <= o1) {  // 32-byte aligned: o0 is a multiple of EMIT_CHUNK
        uint4 *kd = (uint4 *)(a.keys + ob);
        kd[0] = make_uint4(kv[0], kv[1], kv[2], kv[3]);
        kd[1] = make_uint4(kv[4], kv[5], kv[6], kv[7]);
        if (vals) {
            uint4 *vd = (uint4 *)(a.vals + ob);
            vd[0] = make_uint4(vv[0], vv[1], vv[2], vv[3]);
            vd[1] = make_uint4(vv[4], vv[5], vv[6], vv[7]);
        }
        if (a.slot_f) *(uint2 *)(a.slot_f + ob) = make_uint2(0u, 0u);  // output o is the record's slot o
    } else {
#pragma unroll
        for (int i = 0; i < EMIT_ITEMS; i++) {
            if (ob + i >= o1) break;
            a.keys[ob + i] = kv[i];
            if (vals) a.vals[ob + i] = (int32_t)vv[i];
            if (a.slot_f) a.slot_f[ob + i] = 0;
        }
    }
}

// tile_start[t] = first sorted intersection with tile >= t, for t in [0, T]
__global__ __launch_bounds__(256) void tile_ranges_kernel(const uint32_t *keys, const uint32_t *n_ptr,
                                                         const uint32_t *ntiles_ptr, int ntiles_host,
                                                         int32_t *tile_start, int shift) {
    // each thread takes 4 consecutive boundaries k0..k0+3 (one 16-byte load inside the list)
    const uint32_t n = *n_ptr;
    const int ntiles = ntiles_ptr ? (int)*ntiles_ptr : ntiles_host;
    const uint32_t k0 = (blockIdx.x * 256 + threadIdx.x) * 4;
    if (k0 > n) return;
    uint32_t v[4];
    if (k0 + 4 <= n) {
        const uint4 q = *(const uint4 *)(keys + k0);
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) v[j] = k0 + j < n ? keys[k0 + j] : 0u;
    }
    int tp = k0 > 0 ? (int)(keys[k0 - 1] >> shift) : -1;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t k = k0 + j;
        if (k > n) break;
        const int t = k < n ? (int)(v[j] >> shift) : ntiles;
        for (int u = tp + 1; u <= t; u++) tile_start[u] = (int32_t)k;
        tp = t;
    }
}

// View id of each depth-sorted record, the key of one more stable pass that makes the
// record order view-major (each view's records in depth order)
__global__ __launch_bounds__(256) void view_keys_kernel(const int32_t *recs, const uint32_t *n_ptr, int n,
                                                       uint32_t *keys) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < *n_ptr) keys[i] = (uint32_t)(recs[i] / n);
}

// Ranges of a batch sorted by local tile, with each tile's entries grouped by view:
// pair (t, v) owns [range_start[t V + v], range_start[t V + v + 1]), V = num_views.
__global__ __launch_bounds__(256) void tile_ranges_views_kernel(const uint32_t *keys, const int32_t *recs,
                                                               const uint32_t *n_ptr, int n, const BatchTable *tab,
                                                               int32_t *range_start) {
    const uint32_t m = *n_ptr;
    const uint32_t V = (uint32_t)tab->num_views;
    const int total = (int)(tab->max_view_tiles * V);
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k > m) return;
    const int c = k < m ? (int)(keys[k] * V + (uint32_t)(recs[k] / n)) : total;
    const int cp = k > 0 ? (int)(keys[k - 1] * V + (uint32_t)(recs[k - 1] / n)) : -1;
    for (int u = cp + 1; u <= c; u++) range_start[u] = (int32_t)k;
}

__global__ __launch_bounds__(256) void unpack_keys_kernel(const uint32_t *packed, uint32_t n, int shift,
                                                         uint32_t *tiles, int32_t *gid) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const uint32_t v = packed[k];
    tiles[k] = v >> shift;
    gid[k] = (int32_t)(v & ((1u << shift) - 1u));
}

// The 64-bit (tile << 32 | depth bits) view of the sorted intersections
__global__ __launch_bounds__(256) void keys64_kernel(const uint32_t *tiles, const int32_t *gid, const float *depths,
                                                    uint32_t n, uint64_t *out) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) out[k] = ((uint64_t)tiles[k] << 32) | (uint64_t)__float_as_uint(depths[gid[k]]);
}

// ── batched views ────────────────────────────────────────────────────────────

// Projection of gaussian g into every view of the batch: the parameters are read once
// and record v * n + g is written per view, with the same arithmetic as project_kernel
// (the records are bit-identical to one-view projections). Per view, the workgroup's
// intersection and visible-record counts go to block_sums / block_vis[v * nbx + b].
// The SH coefficients of the bands in use by any view (a.sh_u floats per gaussian) are
// staged in LDS once for all views (dynamic LDS: project_sh_lds bytes).
__global__ __launch_bounds__(256) void project_batch_kernel(BatchProjectArgs a) {
    extern __shared__ float shl[];
    __shared__ uint32_t lds4[4];
    __shared__ uint8_t widx[4][64];
    const int g = blockIdx.x * 256 + threadIdx.x;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool live = g < a.n;
    float p[3] = {0.f, 0.f, 0.f}, s[3] = {0.f, 0.f, 0.f}, q[4] = {1.f, 0.f, 0.f, 0.f}, logit = 0.f;
    if (live) {
        p[0] = a.means[3 * g]; p[1] = a.means[3 * g + 1]; p[2] = a.means[3 * g + 2];
        s[0] = exp_cr(a.scales[3 * g]); s[1] = exp_cr(a.scales[3 * g + 1]); s[2] = exp_cr(a.scales[3 * g + 2]);
        q[0] = a.quats[4 * g]; q[1] = a.quats[4 * g + 1]; q[2] = a.quats[4 * g + 2]; q[3] = a.quats[4 * g + 3];
        logit = a.opac[g];
    }
    const float opac = 1.f / (1.f + exp_cr(-logit));
    // staged up to SH degree 3 (sh_u > 0); degree 4 reads its coefficients in place
    const int us = a.sh_u | 1;
    float *wl = shl + (size_t)w * 64 * us;
    if (a.sh_u) stage_sh(live, g - lane, a.dc, a.rest, a.sh_stride, a.sh_u, us, wl, widx[w]);
    const float *mine = wl + lane * us;
    for (int v = 0; v < a.num_views; v++) {
        const msplat_hip_camera &cam = a.tab->views[v].cam;
        const size_t r = (size_t)v * a.n + g;
        uint32_t count = 0;
        if (live) {
            a.radii[r] = 0;
            a.count[r] = 0;
            float pv[3];
            cam_space(cam.viewmat, p, pv);
            if (pv[2] > cam.clip_thresh) {
                float cov[6], c2[3], conic[3], radius;
                covariance3d(s, cam.glob_scale, q, cov);
                const float tfx = 0.5f * (float)cam.width / cam.fx, tfy = 0.5f * (float)cam.height / cam.fy;
                ewa_cov2d(cov, cam.viewmat, cam.fx, cam.fy, tfx, tfy, pv, c2);
                if (conic_radius(c2, conic, radius)) {
                    float px, py;
                    to_pixel(cam.projmat, p, cam.width, cam.height, cam.cx, cam.cy, px, py);
                    const float ax = ceilf(3.f * sqrtf(c2[0])), ay = ceilf(3.f * sqrtf(c2[2]));
                    int x0, y0, x1, y1;
                    tile_rect(px, py, ax, ay, cam.tiles_x, cam.tiles_y, x0, y0, x1, y1);
                    if ((x1 - x0) * (y1 - y0) > 0) {
                        float d[3], col[3];
                        view_direction(p, cam.cam_pos, d);
                        if (a.sh_u) sh_color(cam.degrees_to_use, d, mine, mine + 3, col);
                        else sh_color(cam.degrees_to_use, d, a.dc + 3 * (size_t)g, a.rest + a.sh_stride * g, col);
                        a.splat_a[r] = make_float4(px, py, conic[0], conic[1]);
                        a.splat_b[r] = make_float4(conic[2], opac, col[0], col[1]);
                        a.splat_blue[r] = col[2];
                        a.depths[r] = pv[2];
                        a.rect[r] = make_uint2((uint32_t)x0 | ((uint32_t)x1 << 16),
                                               (uint32_t)y0 | ((uint32_t)y1 << 16));
                        a.radii[r] = (int)radius;
                        count = (uint32_t)((x1 - x0) * (y1 - y0));
                        a.count[r] = count;
                    }
                }
            }
        }
        uint32_t total, vis;
        block_scan_u32(count, total, lds4);
        block_scan_u32(count > 0 ? 1u : 0u, vis, lds4);
        if (threadIdx.x == 0) {
            a.block_sums[(size_t)v * a.nbx + blockIdx.x] = total;
            a.block_vis[(size_t)v * a.nbx + blockIdx.x] = vis;
        }
    }
}

// ── host launchers ───────────────────────────────────────────────────────────

size_t scan_partial_words(uint32_t L) { return (size_t)(L + SCAN_CHUNK - 1) / SCAN_CHUNK + 1; }

void launch_scan_u32(uint32_t *a, uint32_t L, uint32_t *partial, const ScanOut &o, hipStream_t s) {
    if (L <= 2 * SCAN_CHUNK) {  // one launch
        scan_small_kernel<<<1, 256, 0, s>>>(a, L, o);
        return;
    }
    const int nchunks = (int)((L + SCAN_CHUNK - 1) / SCAN_CHUNK);
    if (nchunks > 0) scan_partial_kernel<<<nchunks, 256, 0, s>>>(a, L, partial);
    scan_top_kernel<<<1, 256, 0, s>>>(partial, nchunks, o);
    if (nchunks > 0) scan_apply_kernel<<<nchunks, 256, 0, s>>>(a, L, partial);
}

void launch_scan_u32_pair(uint32_t *a0, const ScanOut &o0, uint32_t *a1, const ScanOut &o1, uint32_t L,
                          uint32_t *partial0, uint32_t *partial1, hipStream_t s) {
    if (L <= 2 * SCAN_CHUNK) {
        scan_small_pair_kernel<<<2, 256, 0, s>>>(a0, o0, a1, o1, L);
        return;
    }
    launch_scan_u32(a0, L, partial0, o0, s);
    launch_scan_u32(a1, L, partial1, o1, s);
}

// LDS of project_kernel's SH staging (four waves of 64 lanes, an odd count of floats per
// lane) for the bands up to `degree`; none when the stored degree is 4 or only the DC is
// in use (read in place)
size_t project_sh_lds(uint32_t degree, uint32_t stored) {
    return stored > 3 || degree == 0 ? 0 : 4 * 64 * (size_t)((3 * sh_bases(degree)) | 1) * sizeof(float);
}

void launch_project(const ProjectArgs &a, hipStream_t s) {
    const int nb = (a.n + 255) / 256;
    if (nb > 0) {
        // a captured step takes the degree in use from device memory: room for every band stored
        if (a.sp) project_kernel<true><<<nb, 256, project_sh_lds(a.cam.degree, a.cam.degree), s>>>(a);
        else project_kernel<false><<<nb, 256, project_sh_lds(a.cam.degrees_to_use, a.cam.degree), s>>>(a);
    }
}

void launch_rect_count(int n, const float4 *splat_a, const float2 *aabb, const int32_t *radii, int tx, int ty,
                       uint2 *rect, uint32_t *count, uint32_t *block_sums, uint32_t *block_vis, hipStream_t s) {
    const int nb = (n + 255) / 256;
    if (nb > 0)
        rect_count_kernel<<<nb, 256, 0, s>>>(n, splat_a, aabb, radii, tx, ty, rect, count, block_sums, block_vis);
}

void launch_scan_blocks(uint32_t *sums, int nblocks, uint64_t *total, uint32_t *total32, hipStream_t s,
                        uint32_t *zero_also) {
    scan_blocks_kernel<<<1, 256, 0, s>>>(sums, nblocks, total, total32, zero_also);
}

void launch_compact(const CompactArgs &a, hipStream_t s) {
    if (a.nbx > 0 && a.num_views > 0) compact_kernel<<<dim3(a.nbx, a.num_views), 256, 0, s>>>(a);
}

// Element blocks of a pass: RS_ITEMS elements from RS_LARGE elements up, RS_ITEMS_SMALL
// below (the histogram words hold the smaller blocks' count either way).
constexpr int RS_ITEMS_SMALL = 1024;
constexpr uint32_t RS_LARGE = 2u << 20;

size_t radix_hist_words(uint32_t n_max) {
    return (size_t)256 * ((n_max + RS_ITEMS_SMALL - 1) / RS_ITEMS_SMALL + 1);
}

// Grid caps of the capacity-sized launches: the workgroups loop over what the device
// count needs.
constexpr uint32_t HIST_GRID = 2048, SCATTER_GRID = 1024, LIST_GRID = 2048;

template <typename K, int ITEMS>
static void radix_pass_items(const K *kin, const int32_t *vin, K *kout, int32_t *vout, const uint32_t *n_dev,
                             uint32_t n_max, int shift, uint32_t mask, uint32_t *hist, uint32_t *digit_total,
                             hipStream_t s) {
    const uint32_t nblocks = (n_max + ITEMS - 1) / ITEMS;
    if (nblocks == 0) return;
    rs_hist_kernel<K, ITEMS><<<std::min(nblocks, HIST_GRID), 256, 0, s>>>(kin, n_dev, shift, mask, hist, (int)nblocks);
    rs_scan_kernel<ITEMS><<<256, 256, 0, s>>>(hist, (int)nblocks, n_dev, digit_total);
    rs_scatter_kernel<K, ITEMS><<<std::min(nblocks, SCATTER_GRID), 256, 0, s>>>(kin, vin, kout, vout, n_dev, shift,
                                                                                mask, hist, (int)nblocks, digit_total);
}

template <typename K>
static void radix_pass(const K *kin, const int32_t *vin, K *kout, int32_t *vout, const uint32_t *n_dev,
                       uint32_t n_max, int shift, uint32_t mask, uint32_t *hist, uint32_t *digit_total, hipStream_t s) {
    if (n_max >= RS_LARGE)
        radix_pass_items<K, RS_ITEMS>(kin, vin, kout, vout, n_dev, n_max, shift, mask, hist, digit_total, s);
    else radix_pass_items<K, RS_ITEMS_SMALL>(kin, vin, kout, vout, n_dev, n_max, shift, mask, hist, digit_total, s);
}

void launch_radix_pass(const uint64_t *kin, const int32_t *vin, uint64_t *kout, int32_t *vout, const uint32_t *n_dev,
                       uint32_t n_max, int shift, uint32_t *hist, uint32_t *digit_total, hipStream_t s, uint32_t mask) {
    radix_pass<uint64_t>(kin, vin, kout, vout, n_dev, n_max, shift, mask, hist, digit_total, s);
}

void launch_radix_pass(const uint32_t *kin, const int32_t *vin, uint32_t *kout, int32_t *vout, const uint32_t *n_dev,
                       uint32_t n_max, int shift, uint32_t *hist, uint32_t *digit_total, hipStream_t s, uint32_t mask) {
    radix_pass<uint32_t>(kin, vin, kout, vout, n_dev, n_max, shift, mask, hist, digit_total, s);
}

void launch_order_sums(const int32_t *list, const uint32_t *nvis_dev, uint32_t nvis_max, const uint32_t *count,
                       uint32_t *sums, uint32_t *lcount, hipStream_t s) {
    const uint32_t nb = (nvis_max + 255) / 256;
    if (nb > 0) order_sums_kernel<<<std::min(nb, LIST_GRID), 256, 0, s>>>(list, nvis_dev, count, sums, nb, lcount);
}

void launch_keygen_offsets(const KeygenArgs &a, hipStream_t s) {
    const uint32_t nb = (a.nvis_max + 255) / 256;
    if (nb > 0) keygen_offsets_kernel<<<std::min(nb, LIST_GRID), 256, 0, s>>>(a);
}

void launch_keygen_emit(const KeygenArgs &a, bool batch, hipStream_t s) {
    const uint32_t ne = (a.m_max + EMIT_CHUNK - 1) / EMIT_CHUNK;
    if (a.nvis_max == 0 || ne == 0) return;
    if (batch) keygen_emit_kernel<true><<<ne, 256, 0, s>>>(a);
    else keygen_emit_kernel<false><<<ne, 256, 0, s>>>(a);
}

void launch_keygen(const KeygenArgs &a, bool batch, hipStream_t s) {
    launch_keygen_offsets(a, s);
    launch_keygen_emit(a, batch, s);
}

void launch_tile_ranges(const uint32_t *keys, const uint32_t *n_dev, uint32_t n_max, const uint32_t *ntiles_dev,
                        int ntiles, int32_t *tile_start, hipStream_t s, int shift) {
    const uint32_t nb = (n_max + 1 + 1023) / 1024;  // 4 boundaries per thread
    tile_ranges_kernel<<<nb, 256, 0, s>>>(keys, n_dev, ntiles_dev, ntiles, tile_start, shift);
}

void launch_unpack_keys(const uint32_t *packed, uint32_t n, int shift, uint32_t *tiles, int32_t *gid, hipStream_t s) {
    if (n) unpack_keys_kernel<<<(n + 255) / 256, 256, 0, s>>>(packed, n, shift, tiles, gid);
}

void launch_keys64(const uint32_t *tiles, const int32_t *gid, const float *depths, uint32_t n, uint64_t *out,
                   hipStream_t s) {
    if (n) keys64_kernel<<<(n + 255) / 256, 256, 0, s>>>(tiles, gid, depths, n, out);
}

void launch_view_keys(const int32_t *recs, const uint32_t *n_dev, uint32_t n_max, int n, uint32_t *keys,
                      hipStream_t s) {
    if (n_max) view_keys_kernel<<<(n_max + 255) / 256, 256, 0, s>>>(recs, n_dev, n, keys);
}

void launch_tile_ranges_views(const uint32_t *keys, const int32_t *recs, const uint32_t *n_dev, uint32_t n_max, int n,
                              const BatchTable *tab, int32_t *range_start, hipStream_t s) {
    tile_ranges_views_kernel<<<(n_max + 1 + 255) / 256, 256, 0, s>>>(keys, recs, n_dev, n, tab, range_start);
}

void launch_batch_project(const BatchProjectArgs &a, hipStream_t s) {
    if (a.nbx > 0) project_batch_kernel<<<a.nbx, 256, 4 * 64 * (size_t)(a.sh_u | 1) * sizeof(float), s>>>(a);
}

}  // namespace msplat
