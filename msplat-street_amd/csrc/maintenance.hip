// Densification (classify, split, duplicate, cull, stable compaction), opacity reset,
// image metrics and RGBA8 packing (core/cpu/msplat_cpu.cpp:1330-1475, 1671-1734).
//
// Densify follows the reference's phases: per-gaussian split/duplicate flags with block
// counts, scans that give every split its child slots N + 2*ord + k and every duplicate
// its slot N + 2*n_split + ord, split children offset by PCG-seeded normals rotated by
// the parent's quaternion, a cull mask over the grown set, and an order-preserving
// compaction of the 18 buffers through a scratch copy.
#include "common.h"
#include "kernels.h"

namespace msplat {

__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t &total, uint32_t *lds4) {
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

__device__ inline float max_exp_scale(const float *s) {
    return fmaxf(fmaxf(exp_cr(s[0]), exp_cr(s[1])), exp_cr(s[2]));
}

// densify_classify (cpu:1354-1369)
__global__ __launch_bounds__(256) void classify_kernel(DensifyArgs a) {
    __shared__ uint32_t lds4[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    uint32_t sp = 0, du = 0;
    if (i < a.n) {
        if (a.vc[i] > 0.0f) {
            const bool high = (a.gn[i] / a.vc[i]) * a.half_max_dim > a.grad_thresh;
            const bool large = max_exp_scale(a.bufs[1] + 3 * (size_t)i) > a.size_thresh;
            const bool split = large || (a.check_screen && a.ms[i] > a.screen_thresh);
            sp = (split && high) ? 1u : 0u;
            du = (!large && high) ? 1u : 0u;
        }
        a.split[i] = (uint8_t)sp;
        a.dup[i] = (uint8_t)du;
    }
    uint32_t ts, td;
    block_excl_scan(sp, ts, lds4);
    block_excl_scan(du, td, lds4);
    if (threadIdx.x == 0) {
        a.split_bs[blockIdx.x] = ts;
        a.dup_bs[blockIdx.x] = td;
    }
}

__device__ inline void copy_row(const DensifyArgs &a, int dst, int src) {
    for (int b = 0; b < 6; b++) {
        const int w = a.width[b];
        float *p = a.bufs[b];
        for (int j = 0; j < w; j++) p[(size_t)dst * w + j] = p[(size_t)src * w + j];
    }
}

__device__ inline void clear_moments(const DensifyArgs &a, int row) {
    for (int b = 6; b < 18; b++) {
        const int w = a.width[b % 6];
        for (int j = 0; j < w; j++) a.bufs[b][(size_t)row * w + j] = 0.f;
    }
}

// densify_append_split (cpu:1387-1415): two children per split, parent shrunk
__global__ __launch_bounds__(256) void split_kernel(DensifyArgs a) {
    __shared__ uint32_t lds4[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t f = i < a.n ? a.split[i] : 0;
    uint32_t t;
    const uint32_t ord = a.split_bs[blockIdx.x] + block_excl_scan(f, t, lds4);
    if (!f) return;
    float *means = a.bufs[0], *scales = a.bufs[1];
    const float *q = a.bufs[2] + 4 * (size_t)i;
    float qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    const float ql = sqrtf(qw * qw + qx * qx + qy * qy + qz * qz);
    qw /= ql; qx /= ql; qy /= ql; qz /= ql;
    const float sx = exp_cr(scales[3 * i]), sy = exp_cr(scales[3 * i + 1]), sz = exp_cr(scales[3 * i + 2]);
    const float shrink = a.log_shrink;
    for (int k = 0; k < 2; k++) {
        const int ch = a.n + 2 * (int)ord + k;
        float nr[3];
        rand_normal3(a.seed, ord * 2u + (uint32_t)k, nr);
        const float r0 = nr[0] * sx, r1 = nr[1] * sy, r2 = nr[2] * sz;
        const float o0 = (1 - 2 * (qy * qy + qz * qz)) * r0 + 2 * (qx * qy - qw * qz) * r1 + 2 * (qx * qz + qw * qy) * r2;
        const float o1 = 2 * (qx * qy + qw * qz) * r0 + (1 - 2 * (qx * qx + qz * qz)) * r1 + 2 * (qy * qz - qw * qx) * r2;
        const float o2 = 2 * (qx * qz - qw * qy) * r0 + 2 * (qy * qz + qw * qx) * r1 + (1 - 2 * (qx * qx + qy * qy)) * r2;
        copy_row(a, ch, i);
        means[3 * (size_t)ch] = means[3 * (size_t)i] + o0;
        means[3 * (size_t)ch + 1] = means[3 * (size_t)i + 1] + o1;
        means[3 * (size_t)ch + 2] = means[3 * (size_t)i + 2] + o2;
        for (int j = 0; j < 3; j++) scales[3 * (size_t)ch + j] = scales[3 * (size_t)i + j] - shrink;
        clear_moments(a, ch);
    }
    for (int j = 0; j < 3; j++) scales[3 * (size_t)i + j] -= shrink;
}

// densify_append_dup (cpu:1417-1430), after every split
__global__ __launch_bounds__(256) void dup_kernel(DensifyArgs a) {
    __shared__ uint32_t lds4[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t f = i < a.n ? a.dup[i] : 0;
    uint32_t t;
    const uint32_t ord = a.dup_bs[blockIdx.x] + block_excl_scan(f, t, lds4);
    if (!f) return;
    const int dst = a.n + 2 * a.n_split + (int)ord;
    copy_row(a, dst, i);
    clear_moments(a, dst);
}

// densify_cull_classify (cpu:1432-1448) over the grown set; keep flags + block counts
__global__ __launch_bounds__(256) void cull_kernel(DensifyArgs a) {
    __shared__ uint32_t lds4[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    uint32_t k = 0;
    if (i < a.n_new) {
        const float o = 1.0f / (1.0f + exp_cr(-a.bufs[5][i]));
        bool cull = o < a.cull_alpha;
        if (i < a.n && a.split[i]) cull = true;
        if (a.check_huge) {
            if (max_exp_scale(a.bufs[1] + 3 * (size_t)i) > a.cull_scale) cull = true;
            if (a.check_screen && i < a.n && a.ms[i] > a.cull_screen) cull = true;
        }
        k = cull ? 0u : 1u;
        a.keep[i] = (uint8_t)k;
    }
    uint32_t t;
    block_excl_scan(k, t, lds4);
    if (threadIdx.x == 0) a.keep_bs[blockIdx.x] = t;
}

// destination row of every kept gaussian (-1 for culled ones)
__global__ __launch_bounds__(256) void keep_rank_kernel(DensifyArgs a) {
    __shared__ uint32_t lds4[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t k = i < a.n_new ? a.keep[i] : 0;
    uint32_t t;
    const uint32_t r = a.keep_bs[blockIdx.x] + block_excl_scan(k, t, lds4);
    if (i < a.n_new) a.dst[i] = k ? (int32_t)r : -1;
}

__global__ __launch_bounds__(256) void compact_scatter_kernel(const float *src, float *scratch, const int32_t *dst,
                                                             int n_new, int width) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (size_t)n_new * width) return;
    const int row = (int)(e / width), j = (int)(e % width);
    const int d = dst[row];
    if (d >= 0) scratch[(size_t)d * width + j] = src[e];
}

__global__ __launch_bounds__(256) void opacity_reset_kernel(float *o, float *m1, float *m2, int n, float reset) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    o[i] = fminf(o[i], reset);
    m1[i] = 0.f;
    m2[i] = 0.f;
}

// RGBA8 by truncation of saturate(c) * 255 (cpu:1723-1734)
__global__ __launch_bounds__(256) void pack_rgba8_kernel(const float *img, uint32_t n, uchar4 *out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float r = fminf(fmaxf(img[3 * i], 0.f), 1.f) * 255.f, g = fminf(fmaxf(img[3 * i + 1], 0.f), 1.f) * 255.f,
                b = fminf(fmaxf(img[3 * i + 2], 0.f), 1.f) * 255.f;
    out[i] = make_uchar4((unsigned char)r, (unsigned char)g, (unsigned char)b, 255);
}

// ── image metrics: clamp-to-edge 11-tap SSIM, L1, PSNR (cpu:1671-1721) ─────────
__device__ constexpr float GM[11] = {0.0010283801f, 0.0075987581f, 0.0360007721f, 0.1093606895f, 0.2130055377f,
                                     0.2660117249f, 0.2130055377f, 0.1093606895f, 0.0360007721f, 0.0075987581f,
                                     0.0010283801f};

__global__ __launch_bounds__(256) void metrics_h_kernel(const float *a, const float *b, int h, int w, float *mom) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= h * w * 3) return;
    const int c = e % 3, x = (e / 3) % w, y = e / (3 * w);
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f;
    for (int k = 0; k < 11; k++) {
        const int sx = min(max(x + k - 5, 0), w - 1);
        const size_t i = ((size_t)y * w + sx) * 3 + c;
        const float g = GM[k], xa = a[i], yv = b[i];
        s0 += g * xa; s1 += g * yv; s2 += g * xa * xa; s3 += g * yv * yv; s4 += g * xa * yv;
    }
    float *o = mom + ((size_t)y * w + x) * 15 + c * 5;
    o[0] = s0; o[1] = s1; o[2] = s2; o[3] = s3; o[4] = s4;
}

__global__ __launch_bounds__(256) void metrics_v_kernel(const float *mom, int h, int w, float *ssim) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= h * w * 3) return;
    const int c = e % 3, x = (e / 3) % w, y = e / (3 * w);
    float s[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < 11; k++) {
        const int sy = min(max(y + k - 5, 0), h - 1);
        const float *o = mom + ((size_t)sy * w + x) * 15 + c * 5;
        for (int q = 0; q < 5; q++) s[q] += GM[k] * o[q];
    }
    const float C1 = 0.0001f, C2 = 0.0009f;
    const float m12 = s[0] * s[1], m1 = s[0] * s[0], m2 = s[1] * s[1];
    const float num = (2.f * m12 + C1) * (2.f * (s[4] - m12) + C2);
    const float den = (m1 + m2 + C1) * ((s[2] - m1) + (s[3] - m2) + C2);
    ssim[e] = num / den;
}

// per-row double sums in (x, channel) order, then rows in order
__global__ __launch_bounds__(64) void metrics_rows_kernel(const float *ssim, const float *a, const float *b, int h,
                                                         int w, double *rows) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    if (y >= h) return;
    double s = 0, l = 0, q = 0;
    for (int x = 0; x < w; x++)
        for (int c = 0; c < 3; c++) {
            const size_t i = ((size_t)y * w + x) * 3 + c;
            s += ssim[i];
            const float d = a[i] - b[i];
            l += fabsf(d);
            q += (double)d * d;
        }
    rows[3 * y] = s;
    rows[3 * y + 1] = l;
    rows[3 * y + 2] = q;
}

__global__ void metrics_total_kernel(const double *rows, int h, int w, double *out) {
    if (threadIdx.x != 0) return;
    double s = 0, l = 0, q = 0;
    for (int y = 0; y < h; y++) {
        s += rows[3 * y];
        l += rows[3 * y + 1];
        q += rows[3 * y + 2];
    }
    const double cnt = (double)h * w * 3, mse = q / cnt;
    out[0] = mse > 0 ? 10.0 * log10(1.0 / mse) : INFINITY;
    out[1] = s / cnt;
    out[2] = l / cnt;
}

// ── launchers ───────────────────────────────────────────────────────────────
static inline int blocks(size_t n) { return (int)((n + 255) / 256); }

void launch_densify_classify(const DensifyArgs &a, hipStream_t s) {
    if (a.n > 0) classify_kernel<<<blocks(a.n), 256, 0, s>>>(a);
}
void launch_densify_append(const DensifyArgs &a, hipStream_t s) {
    if (a.n <= 0) return;
    split_kernel<<<blocks(a.n), 256, 0, s>>>(a);
    dup_kernel<<<blocks(a.n), 256, 0, s>>>(a);
}
void launch_densify_cull(const DensifyArgs &a, hipStream_t s) {
    if (a.n_new > 0) cull_kernel<<<blocks(a.n_new), 256, 0, s>>>(a);
}
void launch_densify_rank(const DensifyArgs &a, hipStream_t s) {
    if (a.n_new > 0) keep_rank_kernel<<<blocks(a.n_new), 256, 0, s>>>(a);
}
void launch_compact_scatter(const float *src, float *scratch, const int32_t *dst, int n_new, int width,
                            hipStream_t s) {
    const size_t total = (size_t)n_new * width;
    if (total) compact_scatter_kernel<<<blocks(total), 256, 0, s>>>(src, scratch, dst, n_new, width);
}
void launch_opacity_reset(float *o, float *m1, float *m2, int n, float reset, hipStream_t s) {
    if (n > 0) opacity_reset_kernel<<<blocks(n), 256, 0, s>>>(o, m1, m2, n, reset);
}
void launch_pack_rgba8(const float *img, uint32_t n, uint8_t *out, hipStream_t s) {
    if (n) pack_rgba8_kernel<<<blocks(n), 256, 0, s>>>(img, n, (uchar4 *)out);
}
void launch_metrics(const float *a, const float *b, int h, int w, float *mom, float *ssim, double *rows,
                    double *out, hipStream_t s) {
    const size_t n = (size_t)h * w * 3;
    metrics_h_kernel<<<blocks(n), 256, 0, s>>>(a, b, h, w, mom);
    metrics_v_kernel<<<blocks(n), 256, 0, s>>>(mom, h, w, ssim);
    metrics_rows_kernel<<<(h + 63) / 64, 64, 0, s>>>(ssim, a, b, h, w, rows);
    metrics_total_kernel<<<1, 64, 0, s>>>(rows, h, w, out);
}

}  // namespace msplat
