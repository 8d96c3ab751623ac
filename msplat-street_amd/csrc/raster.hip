// Per-tile front-to-back alpha compositing (forward) and its backward.
//
// One 256-thread workgroup per 16x16 tile; each wave64 owns an 8x8 quadrant so that a
// wave's pixels are spatially tight (better ballot skipping in the backward). Gaussian
// records of the tile's sorted list are staged through LDS in batches.
//
// Backward without global atomics: per gaussian of a batch, the wave's contributions
// are summed with DPP (wave_sum) only when some lane contributed (ballot), written to a
// per-wave LDS partial, and the four partials are summed in wave order into the
// intersection's gradient slot. Slots are laid out in the pre-sort order of the
// intersections (gaussian-major, tiles row-major), so the per-gaussian reduction reads
// each gaussian's slots contiguously and in the reference's tile order.
#include "common.h"
#include "kernels.h"

namespace msplat {

constexpr int FWD_BATCH = 256;
constexpr int BWD_BATCH = 128;

__device__ inline void tile_pixel(int tile, int tiles_x, int &tx, int &ty, int &px, int &py) {
    tx = tile % tiles_x;
    ty = tile / tiles_x;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    px = tx * TILE + (w & 1) * 8 + (lane & 7);
    py = ty * TILE + (w >> 1) * 8 + (lane >> 3);
}

// nd_rasterize_forward (core/cpu/msplat_cpu.cpp:574-626)
__global__ __launch_bounds__(256) void raster_forward_kernel(RasterArgs a) {
#pragma clang fp contract(fast)
    __shared__ float4 la[FWD_BATCH], lb[FWD_BATCH];
    __shared__ float lc[FWD_BATCH];
    int tx, ty, px, py;
    const int tile = blockIdx.x;
    tile_pixel(tile, a.tiles_x, tx, ty, px, py);
    const bool inside = px < (int)a.width && py < (int)a.height;
    const int s = a.tile_start[tile], e = a.tile_start[tile + 1];
    const float fx = (float)px, fy = (float)py;
    float T = 1.f, cr = 0.f, cg = 0.f, cb = 0.f;
    int last = s - 1;
    bool done = !inside;
    for (int base = s; base < e; base += FWD_BATCH) {
        if (__syncthreads_count(done) == 256) break;
        const int k = base + threadIdx.x;
        if (k < e) {
            const int g = a.gid[k];
            la[threadIdx.x] = a.splat_a[g];
            lb[threadIdx.x] = a.splat_b[g];
            lc[threadIdx.x] = a.splat_blue[g];
        }
        __syncthreads();
        if (!done) {
            const int n = min(FWD_BATCH, e - base);
            for (int j = 0; j < n; j++) {
                const float4 A = la[j], B = lb[j];
                const float dx = A.x - fx, dy = A.y - fy;
                const float sig = 0.5f * (A.z * dx * dx + B.x * dy * dy) + A.w * dx * dy;
                if (sig < 0.f || sig >= SIGMA_MAX) continue;
                const float alpha = fminf(ALPHA_MAX, B.y * __expf(-sig));
                if (alpha < ALPHA_MIN) continue;
                const float nT = T * (1.f - alpha);
                if (nT <= T_MIN) {
                    last = base + j - 1;
                    done = true;
                    break;
                }
                const float vis = alpha * T;
                cr += fmaxf(B.z + 0.5f, 0.f) * vis;
                cg += fmaxf(B.w + 0.5f, 0.f) * vis;
                cb += fmaxf(lc[j] + 0.5f, 0.f) * vis;
                T = nT;
                last = base + j;
            }
        }
    }
    if (inside) {
        const size_t pid = (size_t)py * a.width + px;
        a.final_T[pid] = T;
        a.final_idx[pid] = last;
        a.image[3 * pid] = fminf(fmaxf(a.bg[0] * T + cr, 0.f), 1.f);
        a.image[3 * pid + 1] = fminf(fmaxf(a.bg[1] * T + cg, 0.f), 1.f);
        a.image[3 * pid + 2] = fminf(fmaxf(a.bg[2] * T + cb, 0.f), 1.f);
    }
}

// rasterize_backward (core/cpu/msplat_cpu.cpp:959-1031), slot values per intersection:
// a = (v_x, v_y, v_conic.x, v_conic.y), b = (v_conic.z, v_r, v_g, v_b), c = v_logit.
__global__ __launch_bounds__(256) void raster_backward_kernel(RasterArgs a) {
#pragma clang fp contract(fast)
    __shared__ float4 la[BWD_BATCH], lb[BWD_BATCH];
    __shared__ float lc[BWD_BATCH];
    __shared__ uint32_t lslot[BWD_BATCH];
    __shared__ float part[4][9][BWD_BATCH];
    __shared__ int smax[4];
    int tx, ty, px, py;
    const int tile = blockIdx.x;
    tile_pixel(tile, a.tiles_x, tx, ty, px, py);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool inside = px < (int)a.width && py < (int)a.height;
    const int s = a.tile_start[tile], e = a.tile_start[tile + 1];
    if (s == e) return;
    const size_t pid = (size_t)py * a.width + px;
    float Tf = 0.f, vr = 0.f, vg = 0.f, vb = 0.f;
    int last = s - 1;
    if (inside) {
        Tf = a.final_T[pid];
        last = min(a.final_idx[pid], e - 1);
        vr = a.v_image[3 * pid];
        vg = a.v_image[3 * pid + 1];
        vb = a.v_image[3 * pid + 2];
    }
    // the block's furthest contributing intersection: nothing behind it is visited
    {
        int m = last;
        for (int d = 32; d >= 1; d >>= 1) m = max(m, __shfl_xor(m, d));
        if (lane == 0) smax[w] = m;
    }
    __syncthreads();
    const int kmax = max(max(smax[0], smax[1]), max(smax[2], smax[3]));
    const float tbr = Tf * a.bg[0], tbg = Tf * a.bg[1], tbb = Tf * a.bg[2];
    const float fx = (float)px, fy = (float)py;
    float T = Tf, br = 0.f, bgc = 0.f, bb = 0.f;

    // slots of intersections behind every pixel's last contributor are zero
    for (int k = kmax + 1 + threadIdx.x; k < e; k += 256) {
        const int g = a.gid[k];
        const uint2 r = a.rect[g];
        const int x0 = r.x & 0xFFFF, x1 = r.x >> 16, y0 = r.y & 0xFFFF;
        const uint32_t p = a.offsets[g] + (uint32_t)((ty - y0) * (x1 - x0) + (tx - x0));
        a.slot_a[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        a.slot_b[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        a.slot_c[p] = 0.f;
    }

    for (int end = kmax + 1; end > s; end -= BWD_BATCH) {
        const int start = max(s, end - BWD_BATCH);
        const int n = end - start;
        __syncthreads();
        if (threadIdx.x < n) {
            const int k = start + threadIdx.x;
            const int g = a.gid[k];
            la[threadIdx.x] = a.splat_a[g];
            lb[threadIdx.x] = a.splat_b[g];
            lc[threadIdx.x] = a.splat_blue[g];
            const uint2 r = a.rect[g];
            const int x0 = r.x & 0xFFFF, x1 = r.x >> 16, y0 = r.y & 0xFFFF;
            lslot[threadIdx.x] = a.offsets[g] + (uint32_t)((ty - y0) * (x1 - x0) + (tx - x0));
        }
        for (int i = threadIdx.x; i < 4 * 9 * BWD_BATCH; i += 256) (&part[0][0][0])[i] = 0.f;
        __syncthreads();
        for (int j = n - 1; j >= 0; j--) {
            const int k = start + j;
            float g0 = 0.f, g1 = 0.f, g2 = 0.f, g3 = 0.f, g4 = 0.f, g5 = 0.f, g6 = 0.f, g7 = 0.f, g8 = 0.f;
            bool contrib = false;
            if (k <= last) {
                const float4 A = la[j], B = lb[j];
                const float dx = A.x - fx, dy = A.y - fy;
                const float sig = 0.5f * (A.z * dx * dx + B.x * dy * dy) + A.w * dx * dy;
                if (sig >= 0.f && sig < SIGMA_MAX) {
                    const float alpha = fminf(ALPHA_MAX, B.y * __expf(-sig));
                    if (alpha >= ALPHA_MIN) {
                        contrib = true;
                        const float ra = 1.f / (1.f - alpha);
                        T *= ra;
                        const float fac = alpha * T;
                        const float rr = B.z + 0.5f, rg = B.w + 0.5f, rb = lc[j] + 0.5f;
                        const float cr = fmaxf(rr, 0.f), cg = fmaxf(rg, 0.f), cb = fmaxf(rb, 0.f);
                        const float v_alpha = (cr * T - br * ra - ra * tbr) * vr + (cg * T - bgc * ra - ra * tbg) * vg +
                                              (cb * T - bb * ra - ra * tbb) * vb;
                        br += cr * fac;
                        bgc += cg * fac;
                        bb += cb * fac;
                        g5 = rr >= 0.f ? fac * vr : 0.f;
                        g6 = rg >= 0.f ? fac * vg : 0.f;
                        g7 = rb >= 0.f ? fac * vb : 0.f;
                        if (alpha < ALPHA_MAX) {
                            const float vs = -alpha * v_alpha;
                            g2 = 0.5f * vs * dx * dx;
                            g3 = 0.5f * vs * dx * dy;
                            g4 = 0.5f * vs * dy * dy;
                            g0 = vs * (A.z * dx + A.w * dy);
                            g1 = vs * (A.w * dx + B.x * dy);
                            g8 = -vs * (1.f - B.y);
                        }
                    }
                }
            }
            if (__ballot(contrib) != 0) {
                const float s0 = wave_sum(g0), s1 = wave_sum(g1), s2 = wave_sum(g2), s3 = wave_sum(g3),
                            s4 = wave_sum(g4), s5 = wave_sum(g5), s6 = wave_sum(g6), s7 = wave_sum(g7),
                            s8 = wave_sum(g8);
                if (lane == 0) {
                    part[w][0][j] = s0; part[w][1][j] = s1; part[w][2][j] = s2;
                    part[w][3][j] = s3; part[w][4][j] = s4; part[w][5][j] = s5;
                    part[w][6][j] = s6; part[w][7][j] = s7; part[w][8][j] = s8;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x < n) {
            const int j = threadIdx.x;
            float v[9];
#pragma unroll
            for (int q = 0; q < 9; q++) v[q] = ((part[0][q][j] + part[1][q][j]) + part[2][q][j]) + part[3][q][j];
            const uint32_t p = lslot[j];
            a.slot_a[p] = make_float4(v[0], v[1], v[2], v[3]);
            a.slot_b[p] = make_float4(v[4], v[5], v[6], v[7]);
            a.slot_c[p] = v[8];
        }
    }
}

void launch_raster_forward(const RasterArgs &a, hipStream_t s) {
    const int nt = a.tiles_x * a.tiles_y;
    if (nt > 0) raster_forward_kernel<<<nt, 256, 0, s>>>(a);
}

void launch_raster_backward(const RasterArgs &a, hipStream_t s) {
    const int nt = a.tiles_x * a.tiles_y;
    if (nt > 0) raster_backward_kernel<<<nt, 256, 0, s>>>(a);
}

}  // namespace msplat
