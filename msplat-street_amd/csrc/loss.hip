// L1 + SSIM loss forward and backward (ssim_h_fwd, ssim_fused_v_fwd_h_bwd and
// ssim_v_bwd of core/cpu/msplat_cpu.cpp:669-785), as separable passes over loss blocks
// of 8 rows x 118 columns (128 threads: the block's columns plus 5 on each side):
//   ssim_h     one row of a block per workgroup: horizontal 11-tap moments of every
//              pixel and channel -> planes 0..14;
//   ssim_v     one channel of a block per workgroup: vertical pass, SSIM and its
//              derivative fields (on the block's columns plus the 5-column halo,
//              exchanged through LDS), their horizontal back-convolution -> planes
//              15..23, and the channel's SSIM -> planes 24..26;
//   ssim_vb    one channel of a block per workgroup: vertical back-convolution and
//              dL/d(render);
//   rows       the per-pixel loss (channels' SSIM planes and |target - render| from the
//              images, added in channel order) and its sum per row (for one view in
//              the same launch as ssim_vb);
//   total      the sum over rows.
// In ssim_v and ssim_vb a thread owns one column and slides down the block's rows:
// each input row it loads is added into every output row whose window holds it, so a
// vertical sum costs one load per input row. The planes (108 B per pixel) pass through
// L2 and the memory-side cache between the passes.
// Every sum runs in the reference's tap order with unfused multiply-adds, so the
// gradient and the per-pixel loss are bit-identical to the CPU restatement, and the
// loss total equals the reference's double sum per row (x order) and over rows (y
// order): see row_sum.
#include "common.h"
#include "kernels.h"

namespace msplat {

constexpr int HW = 5;                            // half window
constexpr int LB_THREADS = LOSS_BLOCK_COLS + 2 * HW;
constexpr int LB_IN = LOSS_BLOCK_ROWS + 2 * HW;  // input rows of a block's vertical sums
constexpr int PL_HB = 15, PL_SSIM = 24;  // first planes of the fields and of SSIM
static_assert(LB_THREADS == 128, "two waves per loss block");

__device__ constexpr float GW[11] = {0.0010283801f, 0.0075987581f, 0.0360007721f, 0.1093606895f, 0.2130055377f,
                                     0.2660117249f, 0.2130055377f, 0.1093606895f, 0.0360007721f, 0.0075987581f,
                                     0.0010283801f};

// The target image: the per-step value of a captured step when there is one
__device__ inline const float *target(const LossArgs &a) { return a.sp ? a.sp->gt : a.gt; }

// ssim_h_fwd for row y, columns [X0, X0 + 118): the row's pixels from X0 - 5 to
// X0 + 122 are staged in LDS (zero outside the image), then thread t adds the 11 taps of
// column X0 + t per channel. Plane 5c + q holds moment q (mu_x, mu_y, E[x^2], E[y^2],
// E[xy]; x = target, y = render) of channel c.
__device__ inline void ssim_h_row(const LossArgs &a, int X0, int y) {
    __shared__ float sg[3 * LB_THREADS], sr[3 * LB_THREADS];
    const int W = (int)a.width, H = (int)a.height, t = threadIdx.x;
    if (y >= H) return;
    const size_t P = (size_t)W * H;
    for (int i = t; i < 3 * LB_THREADS; i += LB_THREADS) {
        const int x = X0 - HW + i / 3;
        float gv = 0.f, rv = 0.f;
        if (x >= 0 && x < W) {
            const size_t idx = ((size_t)y * W + x) * 3 + i % 3;
            gv = target(a)[idx];
            rv = a.img[idx];
        }
        sg[i] = gv;
        sr[i] = rv;
    }
    __syncthreads();
    const int x = X0 + t;
    if (t >= LOSS_BLOCK_COLS || x >= W) return;
    for (int c = 0; c < 3; c++) {
        float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f, m4 = 0.f;
#pragma unroll
        for (int k = 0; k < 11; k++) {
            const float gv = sg[3 * (t + k) + c], rv = sr[3 * (t + k) + c], g = GW[k];
            m0 += g * gv;
            m1 += g * rv;
            m2 += g * gv * gv;
            m3 += g * rv * rv;
            m4 += g * gv * rv;
        }
        float *o = a.scratch + (size_t)(5 * c) * P + (size_t)y * W + x;
        o[0] = m0;
        o[P] = m1;
        o[2 * P] = m2;
        o[3 * P] = m3;
        o[4 * P] = m4;
    }
}

// ssim_fused_v_fwd_h_bwd for channel c of the block: thread t owns field column
// xf = X0 - 5 + t. It slides down input rows Y0 - 5 .. Y0 + 12 (rows outside the image
// are skipped, as there), forms SSIM and the fields f1, f2, f3 of its 8 rows (zero
// outside the image), and threads 5..122 then back-convolve the fields of their columns
// from LDS. The same threads store their pixels' SSIM of channel c.
__device__ inline void ssim_v_block(const LossArgs &a, int X0, int Y0, int c) {
    __shared__ float fld[3][LOSS_BLOCK_ROWS][LB_THREADS];
    const int W = (int)a.width, H = (int)a.height, t = threadIdx.x;
    const size_t P = (size_t)W * H;
    const int xf = X0 - HW + t;
    const bool fcol = xf >= 0 && xf < W;
    const bool out = t >= HW && t < HW + LOSS_BLOCK_COLS && xf < W;
    const float C1 = 0.0001f, C2 = 0.0009f;
    float sm[LOSS_BLOCK_ROWS][5];
#pragma unroll
    for (int r = 0; r < LOSS_BLOCK_ROWS; r++)
#pragma unroll
        for (int q = 0; q < 5; q++) sm[r][q] = 0.f;
    if (fcol) {
        const float *m = a.scratch + (size_t)(5 * c) * P + xf;
#pragma unroll
        for (int j = 0; j < LB_IN; j++) {
            const int iy = Y0 - HW + j;
            if (iy < 0 || iy >= H) continue;
            const float *mr = m + (size_t)iy * W;
            const float v0 = mr[0], v1 = mr[P], v2 = mr[2 * P], v3 = mr[3 * P], v4 = mr[4 * P];
#pragma unroll
            for (int r = 0; r < LOSS_BLOCK_ROWS; r++) {
                const int k = j - r;  // the tap of input row iy in output row Y0 + r
                if (k < 0 || k > 10) continue;
                const float g = GW[k];
                sm[r][0] += g * v0;
                sm[r][1] += g * v1;
                sm[r][2] += g * v2;
                sm[r][3] += g * v3;
                sm[r][4] += g * v4;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < LOSS_BLOCK_ROWS; r++) {
        const int y = Y0 + r;
        float f1 = 0.f, f2 = 0.f, f3 = 0.f;
        if (fcol && y < H) {
            const float mx = sm[r][0], my = sm[r][1];
            const float vx = sm[r][2] - mx * mx, vy = sm[r][3] - my * my, cxy = sm[r][4] - mx * my;
            const float A = 2.0f * mx * my + C1, B = 2.0f * cxy + C2;
            const float Cd = mx * mx + my * my + C1, D = vx + vy + C2;
            const float iCD = 1.0f / (Cd * D);
            const float dmu = 2.0f * B * (mx * Cd - A * my) / (Cd * Cd * D);
            const float dvy = -A * B * iCD / D, dcv = 2.0f * A * iCD;
            f1 = dmu - 2.0f * my * dvy - mx * dcv;
            f2 = 2.0f * dvy;
            f3 = dcv;
            if (out) a.scratch[(size_t)(PL_SSIM + c) * P + (size_t)y * W + xf] = (A * B) / (Cd * D);
        }
        fld[0][r][t] = f1;
        fld[1][r][t] = f2;
        fld[2][r][t] = f3;
    }
    __syncthreads();
    if (!out) return;
    float *hb = a.scratch + (size_t)(PL_HB + 3 * c) * P + xf;
    for (int r = 0; r < LOSS_BLOCK_ROWS; r++) {
        const int y = Y0 + r;
        if (y >= H) break;
        float h1 = 0.f, h2 = 0.f, h3 = 0.f;
#pragma unroll
        for (int d = 0; d < 11; d++) {
            const float g = GW[10 - d];
            h1 += g * fld[0][r][t - HW + d];
            h2 += g * fld[1][r][t - HW + d];
            h3 += g * fld[2][r][t - HW + d];
        }
        float *o = hb + (size_t)y * W;
        o[0] = h1;
        o[P] = h2;
        o[2 * P] = h3;
    }
}

// ssim_v_bwd for channel c of the block: thread t owns column X0 + t and slides down the
// back-convolved fields of rows Y0 - 5 .. Y0 + ROWS + 4 (rows outside the image skipped).
template <int ROWS = LOSS_BLOCK_ROWS>
__device__ inline void ssim_vb_block(const LossArgs &a, int X0, int Y0, int c) {
    const int W = (int)a.width, H = (int)a.height, t = threadIdx.x;
    const size_t P = (size_t)W * H;
    const int x = X0 + t;
    if (t >= LOSS_BLOCK_COLS || x >= W) return;
    const float w = a.ssim_weight;
    float cf[ROWS][3];
#pragma unroll
    for (int r = 0; r < ROWS; r++) cf[r][0] = cf[r][1] = cf[r][2] = 0.f;
    const float *hb = a.scratch + (size_t)(PL_HB + 3 * c) * P + x;
#pragma unroll
    for (int j = 0; j < ROWS + 2 * HW; j++) {
        const int iy = Y0 - HW + j;
        if (iy < 0 || iy >= H) continue;
        const float *hr = hb + (size_t)iy * W;
        const float d0 = hr[0], d1 = hr[P], d2 = hr[2 * P];
#pragma unroll
        for (int r = 0; r < ROWS; r++) {
            const int dy = j - r;  // the tap of row iy in output row Y0 + r
            if (dy < 0 || dy > 10) continue;
            const float g = GW[10 - dy];
            cf[r][0] += g * d0;
            cf[r][1] += g * d1;
            cf[r][2] += g * d2;
        }
    }
    for (int r = 0; r < ROWS; r++) {
        const int y = Y0 + r;
        if (y >= H) break;
        const size_t i = ((size_t)y * W + x) * 3 + c;
        const float rv = a.img[i], gv = target(a)[i];
        const float vs = cf[r][0] + rv * cf[r][1] + gv * cf[r][2];
        const float vl = (gv > rv) ? -1.0f : ((gv < rv) ? 1.0f : 0.0f);
        a.v_img[i] = a.inv_n * (-w * vs + (1.0f - w) * vl);
    }
}

// The one-view vertical back-convolution runs taller blocks than the other passes: it
// needs no column halo or LDS, so more rows per thread only cost registers, and its
// 5-row halos are re-read less often.
constexpr int VB_ROWS = 16;

__global__ __launch_bounds__(LB_THREADS) void ssim_h_kernel(LossArgs a) {
    ssim_h_row(a, blockIdx.x * LOSS_BLOCK_COLS, blockIdx.y);
}
__global__ __launch_bounds__(LB_THREADS) void ssim_v_kernel(LossArgs a) {
    ssim_v_block(a, blockIdx.x * LOSS_BLOCK_COLS, blockIdx.y * LOSS_BLOCK_ROWS, blockIdx.z);
}
// The three channel workgroups of a block read and write the same interleaved pixels:
// workgroup ids 8 apart, so they run on one XCD, close in time, and share its L2.
// Block q, channel c is item (q / 8) * 24 + c * 8 + q % 8; items past the blocks exit.
__device__ inline void ssim_vb_item(const LossArgs &a, uint32_t L) {
    const uint32_t rest = L / NUM_XCD;
    const uint32_t c = rest % 3, q = (rest / 3) * NUM_XCD + L % NUM_XCD;
    const uint32_t nbx = (a.width + LOSS_BLOCK_COLS - 1) / LOSS_BLOCK_COLS;
    const uint32_t nby = (a.height + VB_ROWS - 1) / VB_ROWS;
    if (q >= nbx * nby) return;
    ssim_vb_block<VB_ROWS>(a, (int)(q % nbx) * LOSS_BLOCK_COLS, (int)(q / nbx) * VB_ROWS, (int)c);
}

// The reference's per-pixel loss: the channels' SSIM and L1 terms added in channel
// order from zero, then w (1 - ssim / 3) + (1 - w) l1 / 3 in float. The L1 terms come
// from the images (one contiguous 12-byte pixel each), the SSIM terms from their planes.
__device__ inline float pixel_loss(const LossArgs &a, size_t P, size_t p) {
    float s = 0.f, l = 0.f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        s += a.scratch[(size_t)(PL_SSIM + c) * P + p];
        l += fabsf(target(a)[3 * p + c] - a.img[3 * p + c]);
    }
    const float w = a.ssim_weight;
    return w * (1.0f - s / 3.0f) + (1.0f - w) * l / 3.0f;
}

// Exponent e of the ulp of a float (its value is a multiple of 2^e); zero gives
// EXACT_NONE, which constrains nothing.
constexpr int EXACT_NONE = 1 << 20, EXACT_FAIL = -(1 << 20);
__device__ inline int ulp_exp(float v) {
    const uint32_t b = __float_as_uint(v) & 0x7FFFFFFFu;
    if (b == 0) return EXACT_NONE;
    const int e = (int)(b >> 23);
    return e == 0 ? -149 : e - 150;
}

// Whether every partial sum of values that are multiples of 2^emin, with absolute sum
// `abs_sum` (itself summed with rounding, hence the factor-2 margin), is exactly
// representable in double: |partial| <= abs_sum < 2^(emin + 52).
__device__ inline bool sums_exact(int emin, double abs_sum) {
    if (emin == EXACT_NONE) return true;
    if (emin <= EXACT_FAIL / 2 || emin + 51 > 1000) return false;
    return abs_sum <= ldexp(1.0, emin + 51);
}

__device__ inline double wave_sum_d(double v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    return v;
}
__device__ inline int wave_min_i(int v) {
    for (int d = 32; d >= 1; d >>= 1) v = min(v, __shfl_xor(v, d));
    return v;
}

// Row y's loss sum as the reference forms it: a double accumulated over the row's
// pixel losses in x order. When every partial sum is exactly representable
// (sums_exact), each of the reference's additions is exact, so its result is the exact
// sum, which any order of exact additions also gives: the THREADS threads add strided
// values, each wave adds its lanes' sums and thread 0 adds the waves'. Otherwise
// thread 0 recomputes the values and adds them in order (the rare path). Writes the sum,
// the absolute sum and the ulp exponent (EXACT_FAIL when the row was not exact) for the
// total. One workgroup per row.
constexpr int ROW_THREADS = 256;
template <int THREADS = ROW_THREADS>
__device__ inline void row_sum(const LossArgs &a, int y, double *out) {
    __shared__ double ws[THREADS / 64], was[THREADS / 64];
    __shared__ int we[THREADS / 64];
    const int W = (int)a.width, t = threadIdx.x;
    const size_t P = (size_t)W * a.height, base = (size_t)y * W;
    double s = 0.0, as = 0.0;
    int emin = EXACT_NONE;
    for (int x = t; x < W; x += THREADS) {
        const float v = pixel_loss(a, P, base + x);
        s += (double)v;
        as += fabs((double)v);
        emin = min(emin, ulp_exp(v));
    }
    s = wave_sum_d(s);
    as = wave_sum_d(as);
    emin = wave_min_i(emin);
    if ((t & 63) == 0) {
        ws[t >> 6] = s;
        was[t >> 6] = as;
        we[t >> 6] = emin;
    }
    __syncthreads();  // the wave sums
    if (t == 0) {
        s = ws[0];
        as = was[0];
        emin = we[0];
        for (int i = 1; i < THREADS / 64; i++) {
            s += ws[i];
            as += was[i];
            emin = min(emin, we[i]);
        }
        if (!sums_exact(emin, as)) {  // thread 0 adds in x order
            s = 0.0;
            for (int x = 0; x < W; x++)
                s += (double)pixel_loss(a, P, base + x);
            emin = EXACT_FAIL;
        }
        out[3 * y] = s;
        out[3 * y + 1] = as;
        out[3 * y + 2] = (double)emin;
    }
}

// The sum of the row sums as the reference forms it (a double accumulated in y order):
// exact, and so computed as a tree, when every row was exact and the rows together
// pass sums_exact; otherwise one lane adds the row sums in order.
__device__ inline void total_sum(const double *rows, int H, size_t npix, double *sum, float *mean) {
    __shared__ double ws[4], was[4];
    __shared__ int we[4];
    __shared__ int exact;
    const int t = threadIdx.x;
    double s = 0.0, as = 0.0;
    int emin = EXACT_NONE;
    for (int y = t; y < H; y += 256) {
        s += rows[3 * y];
        as += rows[3 * y + 1];
        emin = min(emin, (int)rows[3 * y + 2]);
    }
    s = wave_sum_d(s);
    as = wave_sum_d(as);
    emin = wave_min_i(emin);
    if ((t & 63) == 0) {
        ws[t >> 6] = s;
        was[t >> 6] = as;
        we[t >> 6] = emin;
    }
    __syncthreads();
    if (t == 0) {
        s = ((ws[0] + ws[1]) + ws[2]) + ws[3];
        as = ((was[0] + was[1]) + was[2]) + was[3];
        emin = min(min(we[0], we[1]), min(we[2], we[3]));
        if (!sums_exact(emin, as)) {
            s = 0.0;
            for (int y = 0; y < H; y++) s += rows[3 * y];
        }
        *sum = s;
        *mean = (float)(s / (double)npix);
    }
}

// One view: the row sums and the vertical back-convolution in one launch. They read
// different planes and neither waits for the other, so the latency-bound row sums (the
// first H workgroups) overlap the back-convolution's traffic.
__global__ __launch_bounds__(LB_THREADS) void ssim_vb_rows_kernel(LossArgs a) {
    if (blockIdx.x < a.height) row_sum<LB_THREADS>(a, (int)blockIdx.x, a.row_loss);
    else ssim_vb_item(a, blockIdx.x - a.height);
}

__global__ __launch_bounds__(256) void loss_total_kernel(const double *rows, int H, int W, double *sum, float *mean) {
    total_sum(rows, H, (size_t)W * H, sum, mean);
}

// ── batched views: one launch per stage over every view's tiles / rows ──────────

__device__ inline int view_of_base(const uint32_t *base, int nv, uint32_t i) {
    int lo = 0, hi = nv - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ inline LossArgs view_loss_args(const BatchLossArgs &b, int v) {
    const BatchView &bv = b.tab->views[v];
    const size_t pb = bv.pix_base;
    LossArgs a;
    a.width = bv.cam.width;
    a.height = bv.cam.height;
    a.img = b.img + 3 * pb;
    a.gt = bv.gt;
    a.ssim_weight = b.tab->ssim_weight;
    a.inv_n = bv.inv_n;
    a.v_img = b.v_img + 3 * pb;
    a.scratch = b.scratch + (size_t)LOSS_PLANES * pb;
    a.row_loss = b.row_loss + 3 * (size_t)b.tab->row_base[v];
    a.loss_sum = nullptr;
    a.loss_mean = nullptr;
    return a;
}

// Loss block i of the batch: its view and the block's corner. Grids cover the capacity
// of loss blocks; blocks past the batch's count exit.
__device__ inline bool batch_loss_block(const BatchLossArgs &b, LossArgs &a, int &X0, int &Y0,
                                        uint32_t i = blockIdx.x) {
    if (i >= b.tab->total_loss_blocks) return false;
    const int v = view_of_base(b.tab->loss_base, b.num_views, i);
    a = view_loss_args(b, v);
    const uint32_t local = i - b.tab->loss_base[v];
    const uint32_t nbx = (a.width + LOSS_BLOCK_COLS - 1) / LOSS_BLOCK_COLS;
    X0 = (int)(local % nbx) * LOSS_BLOCK_COLS;
    Y0 = (int)(local / nbx) * LOSS_BLOCK_ROWS;
    return true;
}

__global__ __launch_bounds__(LB_THREADS) void ssim_h_batch_kernel(BatchLossArgs b) {
    LossArgs a;
    int X0, Y0;
    if (batch_loss_block(b, a, X0, Y0)) ssim_h_row(a, X0, Y0 + (int)blockIdx.y);
}
__global__ __launch_bounds__(LB_THREADS) void ssim_v_batch_kernel(BatchLossArgs b) {
    LossArgs a;
    int X0, Y0;
    if (batch_loss_block(b, a, X0, Y0)) ssim_v_block(a, X0, Y0, blockIdx.y);
}
// channel workgroups of a block on one XCD, as in ssim_vb_item
__global__ __launch_bounds__(LB_THREADS) void ssim_vb_batch_kernel(BatchLossArgs b) {
    const uint32_t L = blockIdx.x, rest = L / NUM_XCD;
    const uint32_t c = rest % 3, q = (rest / 3) * NUM_XCD + L % NUM_XCD;
    LossArgs a;
    int X0, Y0;
    if (batch_loss_block(b, a, X0, Y0, q)) ssim_vb_block(a, X0, Y0, (int)c);
}

__global__ __launch_bounds__(ROW_THREADS) void loss_rows_batch_kernel(BatchLossArgs b) {
    const uint32_t q = blockIdx.x;
    if (q >= b.tab->total_rows) return;
    const int v = view_of_base(b.tab->row_base, b.num_views, q);
    const LossArgs a = view_loss_args(b, v);
    row_sum(a, (int)(q - b.tab->row_base[v]), a.row_loss);
}

// one workgroup per view: the sum of its rows, and the mean
__global__ __launch_bounds__(256) void loss_total_batch_kernel(BatchLossArgs b) {
    const int v = blockIdx.x;
    const BatchView &bv = b.tab->views[v];
    total_sum(b.row_loss + 3 * (size_t)b.tab->row_base[v], (int)bv.cam.height,
              (size_t)bv.cam.width * bv.cam.height, &b.sc->loss_sum[v], &b.sc->loss_mean[v]);
}

void launch_batch_loss(const BatchLossArgs &b, hipStream_t s) {
    if (b.num_views <= 0 || b.max_loss_blocks == 0) return;
    ssim_h_batch_kernel<<<dim3(b.max_loss_blocks, LOSS_BLOCK_ROWS), LB_THREADS, 0, s>>>(b);
    ssim_v_batch_kernel<<<dim3(b.max_loss_blocks, 3), LB_THREADS, 0, s>>>(b);
    ssim_vb_batch_kernel<<<3 * ((b.max_loss_blocks + NUM_XCD - 1) / NUM_XCD) * NUM_XCD, LB_THREADS, 0, s>>>(b);
    loss_rows_batch_kernel<<<b.max_rows, ROW_THREADS, 0, s>>>(b);
    loss_total_batch_kernel<<<b.num_views, 256, 0, s>>>(b);
}

void launch_loss(const LossArgs &a, hipStream_t s) {
    if (a.width == 0 || a.height == 0) return;
    const uint32_t nbx = (a.width + LOSS_BLOCK_COLS - 1) / LOSS_BLOCK_COLS;
    const uint32_t nby = (a.height + LOSS_BLOCK_ROWS - 1) / LOSS_BLOCK_ROWS;
    ssim_h_kernel<<<dim3(nbx, a.height), LB_THREADS, 0, s>>>(a);
    ssim_v_kernel<<<dim3(nbx, nby, 3), LB_THREADS, 0, s>>>(a);
    const uint32_t nvb = nbx * ((a.height + VB_ROWS - 1) / VB_ROWS);
    ssim_vb_rows_kernel<<<a.height + 3 * ((nvb + NUM_XCD - 1) / NUM_XCD) * NUM_XCD, LB_THREADS, 0, s>>>(a);
    loss_total_kernel<<<1, 256, 0, s>>>(a.row_loss, a.height, a.width, a.loss_sum, a.loss_mean);
}

}  // namespace msplat
