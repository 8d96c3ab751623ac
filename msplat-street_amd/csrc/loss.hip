// Fused L1 + SSIM loss forward and backward (ssim_h_fwd, ssim_fused_v_fwd_h_bwd and
// ssim_v_bwd of core/cpu/msplat_cpu.cpp:669-785 in one pass over the image).
//
// One 256-thread workgroup per 32x16 output tile. Per channel, the tile's input window
// (tile + 10 px halo, zero outside the image) is staged in LDS; horizontal 11-tap
// moments, the vertical pass with SSIM and its derivative fields (f1, f2, f3 on the
// tile + 5 px), and both back-convolutions are computed from LDS. HBM traffic is the
// minimum: render and target read once, dL/d(render) and a per-pixel loss written once.
// Every sum runs in the reference's tap order with unfused multiply-adds, so the
// gradient and the per-pixel loss are bit-identical to the CPU restatement; the loss
// total is accumulated in double per row (x order) and over rows (y order) as there.
#include "common.h"
#include "kernels.h"

namespace msplat {

constexpr int LT_W = 32, LT_H = 16;
constexpr int IN_W = LT_W + 20, IN_H = LT_H + 20;  // input window
constexpr int MH_W = LT_W + 10;                     // moments / field columns
constexpr int F_H = LT_H + 10;                      // field rows

__device__ constexpr float GW[11] = {0.0010283801f, 0.0075987581f, 0.0360007721f, 0.1093606895f, 0.2130055377f,
                                     0.2660117249f, 0.2130055377f, 0.1093606895f, 0.0360007721f, 0.0075987581f,
                                     0.0010283801f};

// The loss of the 32x16 output tile at (X0, Y0); every thread of the workgroup calls it.
__device__ inline void loss_tile(const LossArgs &a, int X0, int Y0) {
    __shared__ float in_g[IN_H][IN_W], in_r[IN_H][IN_W];
    __shared__ float mh[5][IN_H][MH_W];  // reused for the horizontal back-convolution
    __shared__ float fld[3][F_H][MH_W];
    __shared__ float acc_s[LT_H * LT_W], acc_l[LT_H * LT_W];
    float(*hb)[F_H][LT_W] = reinterpret_cast<float(*)[F_H][LT_W]>(&mh[0][0][0]);

    const int W = (int)a.width, H = (int)a.height;
    const int tid = threadIdx.x;
    const float C1 = 0.0001f, C2 = 0.0009f;
    const float w = a.ssim_weight;

    for (int i = tid; i < LT_H * LT_W; i += 256) {
        acc_s[i] = 0.f;
        acc_l[i] = 0.f;
    }
    for (int c = 0; c < 3; c++) {
        // input window of channel c
        for (int i = tid; i < IN_H * IN_W; i += 256) {
            const int r = i / IN_W, col = i % IN_W;
            const int y = Y0 - 10 + r, x = X0 - 10 + col;
            float gv = 0.f, rv = 0.f;
            if (x >= 0 && x < W && y >= 0 && y < H) {
                const size_t idx = ((size_t)y * W + x) * 3 + c;
                gv = a.gt[idx];
                rv = a.img[idx];
            }
            in_g[r][col] = gv;
            in_r[r][col] = rv;
        }
        __syncthreads();
        // horizontal moments mu_x, mu_y, E[x^2], E[y^2], E[xy] (x = target, y = render)
        for (int i = tid; i < IN_H * MH_W; i += 256) {
            const int r = i / MH_W, col = i % MH_W;
            float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f, m4 = 0.f;
#pragma unroll
            for (int k = 0; k < 11; k++) {
                const float gv = in_g[r][col + k], rv = in_r[r][col + k], g = GW[k];
                m0 += g * gv;
                m1 += g * rv;
                m2 += g * gv * gv;
                m3 += g * rv * rv;
                m4 += g * gv * rv;
            }
            mh[0][r][col] = m0; mh[1][r][col] = m1; mh[2][r][col] = m2; mh[3][r][col] = m3; mh[4][r][col] = m4;
        }
        __syncthreads();
        // vertical pass, SSIM, derivative fields
        for (int i = tid; i < F_H * MH_W; i += 256) {
            const int fr = i / MH_W, col = i % MH_W;
            const int y = Y0 - 5 + fr, x = X0 - 5 + col;
            float f1 = 0.f, f2 = 0.f, f3 = 0.f;
            if (x >= 0 && x < W && y >= 0 && y < H) {
                float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f;
#pragma unroll
                for (int k = 0; k < 11; k++) {
                    const int sy = y - 5 + k;
                    if (sy < 0 || sy >= H) continue;
                    const float g = GW[k];
                    s0 += g * mh[0][fr + k][col];
                    s1 += g * mh[1][fr + k][col];
                    s2 += g * mh[2][fr + k][col];
                    s3 += g * mh[3][fr + k][col];
                    s4 += g * mh[4][fr + k][col];
                }
                const float mx = s0, my = s1;
                const float vx = s2 - mx * mx, vy = s3 - my * my, cxy = s4 - mx * my;
                const float A = 2.0f * mx * my + C1, B = 2.0f * cxy + C2;
                const float Cd = mx * mx + my * my + C1, D = vx + vy + C2;
                const float iCD = 1.0f / (Cd * D);
                const float dmu = 2.0f * B * (mx * Cd - A * my) / (Cd * Cd * D);
                const float dvy = -A * B * iCD / D, dcv = 2.0f * A * iCD;
                f1 = dmu - 2.0f * my * dvy - mx * dcv;
                f2 = 2.0f * dvy;
                f3 = dcv;
                const int oy = fr - 5, ox = col - 5;
                if (oy >= 0 && oy < LT_H && ox >= 0 && ox < LT_W) {
                    const int o = oy * LT_W + ox;
                    acc_s[o] += (A * B) / (Cd * D);
                    acc_l[o] += fabsf(in_g[fr + 5][col + 5] - in_r[fr + 5][col + 5]);
                }
            }
            fld[0][fr][col] = f1; fld[1][fr][col] = f2; fld[2][fr][col] = f3;
        }
        __syncthreads();
        // horizontal back-convolution of the fields
        for (int i = tid; i < F_H * LT_W; i += 256) {
            const int fr = i / LT_W, ox = i % LT_W;
            float h1 = 0.f, h2 = 0.f, h3 = 0.f;
#pragma unroll
            for (int d = 0; d < 11; d++) {
                const float g = GW[10 - d];
                h1 += g * fld[0][fr][ox + d];
                h2 += g * fld[1][fr][ox + d];
                h3 += g * fld[2][fr][ox + d];
            }
            hb[0][fr][ox] = h1; hb[1][fr][ox] = h2; hb[2][fr][ox] = h3;
        }
        __syncthreads();
        // vertical back-convolution and the image gradient
        for (int i = tid; i < LT_H * LT_W; i += 256) {
            const int oy = i / LT_W, ox = i % LT_W;
            const int y = Y0 + oy, x = X0 + ox;
            if (x >= W || y >= H) continue;
            float a1 = 0.f, a2 = 0.f, a3 = 0.f;
#pragma unroll
            for (int d = 0; d < 11; d++) {
                const int sy = y - 5 + d;
                if (sy < 0 || sy >= H) continue;
                const float g = GW[10 - d];
                a1 += g * hb[0][oy + d][ox];
                a2 += g * hb[1][oy + d][ox];
                a3 += g * hb[2][oy + d][ox];
            }
            const float rv = in_r[oy + 10][ox + 10], gv = in_g[oy + 10][ox + 10];
            const float vs = a1 + rv * a2 + gv * a3;
            const float vl = (gv > rv) ? -1.0f : ((gv < rv) ? 1.0f : 0.0f);
            a.v_img[((size_t)y * W + x) * 3 + c] = a.inv_n * (-w * vs + (1.0f - w) * vl);
        }
        __syncthreads();
    }
    for (int i = tid; i < LT_H * LT_W; i += 256) {
        const int y = Y0 + i / LT_W, x = X0 + i % LT_W;
        if (x < W && y < H) a.pix_loss[(size_t)y * W + x] = w * (1.0f - acc_s[i] / 3.0f) + (1.0f - w) * acc_l[i] / 3.0f;
    }
}

__global__ __launch_bounds__(256) void loss_kernel(LossArgs a) { loss_tile(a, blockIdx.x * LT_W, blockIdx.y * LT_H); }

// Per-row double sums of the per-pixel loss in x order: the row is staged in LDS with
// coalesced loads, then one lane adds it up sequentially (the reference's order).
constexpr int ROW_MAX = 4096;

__global__ __launch_bounds__(256) void loss_rows_kernel(const float *pix, int W, int H, double *rows) {
    __shared__ float r[ROW_MAX];
    const int y = blockIdx.x;
    const float *p = pix + (size_t)y * W;
    double s = 0.0;
    for (int base = 0; base < W; base += ROW_MAX) {
        const int n = min(ROW_MAX, W - base);
        __syncthreads();
        for (int x = threadIdx.x; x < n; x += 256) r[x] = p[base + x];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int x = 0; x < n; x++) s += (double)r[x];
    }
    if (threadIdx.x == 0) rows[y] = s;
}

// Sum of the row sums in y order (staged in LDS, one lane adds)
__global__ __launch_bounds__(256) void loss_total_kernel(const double *rows, int H, int W, double *sum, float *mean) {
    __shared__ double r[2048];
    double s = 0.0;
    for (int base = 0; base < H; base += 2048) {
        const int n = min(2048, H - base);
        __syncthreads();
        for (int y = threadIdx.x; y < n; y += 256) r[y] = rows[base + y];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int y = 0; y < n; y++) s += r[y];
    }
    if (threadIdx.x == 0) {
        *sum = s;
        *mean = (float)(s / (double)((size_t)W * H));
    }
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
    a.pix_loss = b.pix_loss + pb;
    a.row_loss = nullptr;
    a.loss_sum = nullptr;
    a.loss_mean = nullptr;
    return a;
}

// grid = the capacity of loss tiles; workgroups past the batch's count exit
__global__ __launch_bounds__(256) void loss_batch_kernel(BatchLossArgs b) {
    const uint32_t i = blockIdx.x;
    if (i >= b.tab->total_loss_blocks) return;
    const int v = view_of_base(b.tab->loss_base, b.num_views, i);
    const LossArgs a = view_loss_args(b, v);
    const uint32_t local = i - b.tab->loss_base[v];
    const uint32_t nbx = (a.width + LT_W - 1) / LT_W;
    loss_tile(a, (int)(local % nbx) * LT_W, (int)(local / nbx) * LT_H);
}

__global__ __launch_bounds__(256) void loss_rows_batch_kernel(BatchLossArgs b) {
    const uint32_t q = blockIdx.x;
    if (q >= b.tab->total_rows) return;
    const int v = view_of_base(b.tab->row_base, b.num_views, q);
    const BatchView &bv = b.tab->views[v];
    const int W = (int)bv.cam.width;
    const int y = (int)(q - b.tab->row_base[v]);
    __shared__ float r[ROW_MAX];
    const float *p = b.pix_loss + bv.pix_base + (size_t)y * W;
    double s = 0.0;
    for (int base = 0; base < W; base += ROW_MAX) {
        const int n = min(ROW_MAX, W - base);
        __syncthreads();
        for (int x = threadIdx.x; x < n; x += 256) r[x] = p[base + x];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int x = 0; x < n; x++) s += (double)r[x];
    }
    if (threadIdx.x == 0) b.row_loss[q] = s;
}

// one workgroup per view: the sum of its rows in y order, and the mean
__global__ __launch_bounds__(256) void loss_total_batch_kernel(BatchLossArgs b) {
    const int v = blockIdx.x;
    const BatchView &bv = b.tab->views[v];
    const int H = (int)bv.cam.height, W = (int)bv.cam.width;
    const double *rows = b.row_loss + b.tab->row_base[v];
    __shared__ double r[2048];
    double s = 0.0;
    for (int base = 0; base < H; base += 2048) {
        const int n = min(2048, H - base);
        __syncthreads();
        for (int y = threadIdx.x; y < n; y += 256) r[y] = rows[base + y];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int y = 0; y < n; y++) s += r[y];
    }
    if (threadIdx.x == 0) {
        b.sc->loss_sum[v] = s;
        b.sc->loss_mean[v] = (float)(s / (double)((size_t)W * H));
    }
}

void launch_batch_loss(const BatchLossArgs &b, hipStream_t s) {
    if (b.num_views <= 0) return;
    loss_batch_kernel<<<b.max_loss_blocks, 256, 0, s>>>(b);
    loss_rows_batch_kernel<<<b.max_rows, 256, 0, s>>>(b);
    loss_total_batch_kernel<<<b.num_views, 256, 0, s>>>(b);
}

void launch_loss(const LossArgs &a, hipStream_t s) {
    dim3 grid((a.width + LT_W - 1) / LT_W, (a.height + LT_H - 1) / LT_H);
    loss_kernel<<<grid, 256, 0, s>>>(a);
    loss_rows_kernel<<<a.height, 256, 0, s>>>(a.pix_loss, a.width, a.height, a.row_loss);
    loss_total_kernel<<<1, 256, 0, s>>>(a.row_loss, a.height, a.width, a.loss_sum, a.loss_mean);
}

}  // namespace msplat
