// Prior-guided training extras (PriorStep, core/metal/bindings.h:43-77): the learned sky
// behind the gaussians, the exposure transform of the photometric loss, the depth / sky /
// fill prior losses with their gradients, the Adam steps of the sky texture and of the
// exposure rows, and the needle cap (core/cpu/msplat_cpu.cpp:327-402, 787-950,
// 1219-1309). One thread per pixel, texel or gaussian. Where the reference sums in
// double, so does this file, in fixed order. The sky gradient gathers many pixels into
// each texel: the taps are added as 64-bit fixed-point integers (2^-48 units), whose
// atomic sums do not depend on their order, so every run gives the same texture.
#include "common.h"
#include "kernels.h"

namespace msplat {

namespace {

constexpr float kPi = 3.14159265358979323846f;

// world direction of the ray through pixel (px, py) (pixel_world_dir, cpu:353-359)
__device__ inline void pixel_dir(const SkyArgs &k, float px, float py, float d[3]) {
    const float dx = (px + 0.5f - k.cx) / k.fx, dy = (py + 0.5f - k.cy) / k.fy;
    d[0] = k.vm[0] * dx + k.vm[4] * dy + k.vm[8];
    d[1] = k.vm[1] * dx + k.vm[5] * dy + k.vm[9];
    d[2] = k.vm[2] * dx + k.vm[6] * dy + k.vm[10];
}

// the equirect taps of that ray (sky_uv + sky_taps, cpu:361-395): azimuth around `up`
// from e1 towards e2 wraps in u, elevation clamps in v
__device__ inline void sky_taps(const SkyArgs &k, float px, float py, uint32_t idx[4], float wt[4]) {
    float d[3];
    pixel_dir(k, px, py, d);
    const float inv = 1.f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    d[0] *= inv; d[1] *= inv; d[2] *= inv;
    const float du = d[0] * k.up[0] + d[1] * k.up[1] + d[2] * k.up[2];
    const float d1 = d[0] * k.e1[0] + d[1] * k.e1[1] + d[2] * k.e1[2];
    const float d2 = d[0] * k.e2[0] + d[1] * k.e2[1] + d[2] * k.e2[2];
    const float el = asinf(fminf(fmaxf(du, -1.f), 1.f)), az = atan2f(d2, d1);
    const float u = (az * (0.5f / kPi) + 0.5f) * (float)k.w - 0.5f;
    const float v = (0.5f - el / kPi) * (float)k.h - 0.5f;
    const float u0f = floorf(u), fu = u - u0f;
    int u0 = (int)u0f % (int)k.w;
    if (u0 < 0) u0 += (int)k.w;
    const int u1 = (u0 + 1) % (int)k.w;
    const float vc = fminf(fmaxf(v, 0.f), (float)k.h - 1.f);
    const int v0 = (int)vc, v1 = min(v0 + 1, (int)k.h - 1);
    const float fv = vc - (float)v0;
    idx[0] = (uint32_t)(v0 * (int)k.w + u0); wt[0] = (1.f - fu) * (1.f - fv);
    idx[1] = (uint32_t)(v0 * (int)k.w + u1); wt[1] = fu * (1.f - fv);
    idx[2] = (uint32_t)(v1 * (int)k.w + u0); wt[2] = (1.f - fu) * fv;
    idx[3] = (uint32_t)(v1 * (int)k.w + u1); wt[3] = fu * fv;
}

// The background behind every pixel: the sky's bilinear sample, or the constant
__global__ __launch_bounds__(256) void sky_background_kernel(SkyArgs k, uint32_t W, uint32_t H, float b0, float b1,
                                                             float b2, float *bg_img) {
    const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    float c[3] = {b0, b1, b2};
    if (k.tex) {
        uint32_t idx[4];
        float wt[4];
        sky_taps(k, (float)x, (float)y, idx, wt);
        c[0] = c[1] = c[2] = 0.f;
        for (int t = 0; t < 4; t++)
            for (int ch = 0; ch < 3; ch++) c[ch] = k.tex[3 * (size_t)idx[t] + ch] * wt[t] + c[ch];
    }
    const size_t p = (size_t)y * W + x;
    bg_img[3 * p] = c[0];
    bg_img[3 * p + 1] = c[1];
    bg_img[3 * p + 2] = c[2];
}

// exposure_apply (cpu:887-896): out = M c + o, with M the row's first nine values
__global__ __launch_bounds__(256) void exposure_apply_kernel(const float *img, const float *P, size_t n, float *out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float c0 = img[3 * i], c1 = img[3 * i + 1], c2 = img[3 * i + 2];
    out[3 * i] = P[0] * c0 + P[1] * c1 + P[2] * c2 + P[9];
    out[3 * i + 1] = P[3] * c0 + P[4] * c1 + P[5] * c2 + P[10];
    out[3 * i + 2] = P[6] * c0 + P[7] * c1 + P[8] * c2 + P[11];
}

// sample_prior_depth (cpu:790-806): bilinear, nearest across a depth edge or invalid sample
__device__ inline float prior_depth_at(const float *d, float x, float y, uint32_t w, uint32_t h) {
    x = fminf(fmaxf(x, 0.f), (float)w - 1.f);
    y = fminf(fmaxf(y, 0.f), (float)h - 1.f);
    const uint32_t x0 = (uint32_t)x, y0 = (uint32_t)y, x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
    const float fx = x - (float)x0, fy = y - (float)y0;
    const float d00 = d[y0 * w + x0], d01 = d[y0 * w + x1], d10 = d[y1 * w + x0], d11 = d[y1 * w + x1];
    const float lo = fminf(fminf(d00, d01), fminf(d10, d11)), hi = fmaxf(fmaxf(d00, d01), fmaxf(d10, d11));
    if (lo <= 0.f || hi > 1.05f * lo) return d[((fy < 0.5f) ? y0 : y1) * w + ((fx < 0.5f) ? x0 : x1)];
    const float top = d00 + (d01 - d00) * fx, bottom = d10 + (d11 - d10) * fx;
    return top + (bottom - top) * fy;
}

// A double moved across lanes by DPP (both halves with the same control)
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline double dpp_d(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROW_MASK, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROW_MASK, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The wave's total of a double in lane 63, in a fixed order (the steps of a float wave sum)
__device__ inline double wave_total_d(double v) {
    v += dpp_d<0xb1>(v);
    v += dpp_d<0x4e>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    v += dpp_d<0x142, 0xa>(v);
    v += dpp_d<0x143, 0xc>(v);
    return v;
}

constexpr int PL_CHUNK = 1024;  // pixels of a row per round (4 per thread)

// prior_loss (cpu:809-884), one workgroup per row. The row's three sums of the pixels'
// float terms are taken in double in a fixed order: each thread's pixels in order, each
// wave's lanes by a DPP reduction, the four waves in order (the reference adds the row
// in pixel order; the two orders round differently by about 1e-16 relative).
__global__ __launch_bounds__(256) void prior_loss_kernel(PriorLossArgs a) {
    __shared__ double wsum[4][3];
    double tsum[3] = {0.0, 0.0, 0.0};
    const uint32_t gy = blockIdx.x, W = a.width, H = a.height;
    const bool has_prior = a.prior_depth != nullptr;
    const uint32_t pw = has_prior ? a.pw : 1u, ph = has_prior ? a.ph : 1u;
    const float inv_npix = 1.0f / (float)(W * H);
    for (uint32_t x0 = 0; x0 < W; x0 += PL_CHUNK) {
        for (int u = 0; u < PL_CHUNK / 256; u++) {
            const uint32_t j = u * 256 + threadIdx.x, gx = x0 + j;
            float ld = 0.f, ls = 0.f, lf = 0.f;
            if (gx < W) {
                const size_t pix = (size_t)gy * W + gx;
                const float sx = ((float)gx + 0.5f) * (float)pw / (float)W;
                const float sy = ((float)gy + 0.5f) * (float)ph / (float)H;
                float conf = 0.f;
                bool is_sky = false, keep = true;
                if (has_prior) {
                    const uint32_t q = min((uint32_t)sy, ph - 1) * pw + min((uint32_t)sx, pw - 1);
                    const uchar4 pa = reinterpret_cast<const uchar4 *>(a.prior_aux)[q];
                    conf = (float)pa.x * (1.f / 255.f);
                    is_sky = a.has_sky && pa.y >= 128;
                    keep = pa.z >= 128;
                }
                float v_d = 0.f, v_a = 0.f;
                if (!keep) {
                    if (a.mask_photometric) a.v_photo[3 * pix] = a.v_photo[3 * pix + 1] = a.v_photo[3 * pix + 2] = 0.f;
                } else {
                    const float A = 1.f - a.final_T[pix];
                    if (is_sky) {
                        ls = a.sky_weight * A;
                        v_a = a.sky_weight * inv_npix;
                    } else {
                        if (a.has_sky && a.fill_weight > 0.f) {
                            lf = a.fill_weight * (1.f - A);
                            v_a -= a.fill_weight * inv_npix;
                        }
                        const float D = a.out_depth[pix];
                        if (a.has_depth && a.depth_weight > 0.f && conf > 0.f && A > a.min_alpha && D > 0.f) {
                            const float dp = prior_depth_at(a.prior_depth, sx - 0.5f, sy - 0.5f, pw, ph);
                            if (dp > 0.f) {
                                const float r = logf(D / A) - logf(dp), ar = fabsf(r);
                                float dr;
                                if (ar < a.huber) {
                                    ld = a.depth_weight * conf * 0.5f * r * r / a.huber;
                                    dr = r / a.huber;
                                } else {
                                    ld = a.depth_weight * conf * (ar - 0.5f * a.huber);
                                    dr = r > 0.f ? 1.f : (r < 0.f ? -1.f : 0.f);
                                }
                                const float sc = a.depth_weight * conf * dr * inv_npix;
                                v_d = sc / D;
                                v_a -= sc / A;
                            }
                        }
                    }
                }
                a.v_aux[2 * pix] = v_d;
                a.v_aux[2 * pix + 1] = v_a;
            }
            tsum[0] += (double)ld;
            tsum[1] += (double)ls;
            tsum[2] += (double)lf;
        }
    }
    for (int k = 0; k < 3; k++) {
        const double t = wave_total_d(tsum[k]);
        if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6][k] = t;
    }
    __syncthreads();
    if (threadIdx.x < 3)
        a.rows[3 * (size_t)gy + threadIdx.x] =
            ((wsum[0][threadIdx.x] + wsum[1][threadIdx.x]) + wsum[2][threadIdx.x]) + wsum[3][threadIdx.x];
}

// the terms' sums over the rows, scaled by 1 / (W H) (the slot after them is 0), in a
// fixed order: each thread's rows y = t, t + 256, ... in order, each wave by a DPP
// reduction, the four waves in order
__global__ __launch_bounds__(256) void prior_terms_kernel(const double *rows, uint32_t H, float inv_npix, float *lt) {
    __shared__ double wsum[4][3];
    double t[3] = {0.0, 0.0, 0.0};
    for (uint32_t y = threadIdx.x; y < H; y += 256)
        for (int k = 0; k < 3; k++) t[k] += rows[3 * (size_t)y + k];
    for (int k = 0; k < 3; k++) {
        const double w = wave_total_d(t[k]);
        if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6][k] = w;
    }
    __syncthreads();
    if (threadIdx.x < 3)
        lt[threadIdx.x] = (float)((((wsum[0][threadIdx.x] + wsum[1][threadIdx.x]) + wsum[2][threadIdx.x]) +
                                   wsum[3][threadIdx.x]) * (double)inv_npix);
    if (threadIdx.x == 3) lt[3] = 0.f;
}

constexpr int EX_ITEMS = 4096;  // pixels per exposure block
constexpr int EX_THREADS = 1024;

// exposure_backward (cpu:899-913): v_img = M^T v_adj per pixel, and the block's sums of
// the 12 gradient terms in double: each thread's pixels in order, then each wave's lanes
// by a fixed DPP reduction, then the 16 waves in order
__global__ __launch_bounds__(EX_THREADS) void exposure_backward_kernel(const float *img, const float *v_adj, size_t n,
                                                                      const float *P, float *v_img, double *part) {
    __shared__ double red[EX_THREADS / 64][12];
    double g[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const size_t base = (size_t)blockIdx.x * EX_ITEMS;
    for (int u = 0; u < EX_ITEMS / EX_THREADS; u++) {
        const size_t i = base + (size_t)u * EX_THREADS + threadIdx.x;
        if (i >= n) break;
        const float c0 = img[3 * i], c1 = img[3 * i + 1], c2 = img[3 * i + 2];
        const float vx = v_adj[3 * i], vy = v_adj[3 * i + 1], vz = v_adj[3 * i + 2];
        v_img[3 * i] = P[0] * vx + P[3] * vy + P[6] * vz;
        v_img[3 * i + 1] = P[1] * vx + P[4] * vy + P[7] * vz;
        v_img[3 * i + 2] = P[2] * vx + P[5] * vy + P[8] * vz;
        g[0] += vx * c0; g[1] += vx * c1; g[2] += vx * c2;
        g[3] += vy * c0; g[4] += vy * c1; g[5] += vy * c2;
        g[6] += vz * c0; g[7] += vz * c1; g[8] += vz * c2;
        g[9] += vx; g[10] += vy; g[11] += vz;
    }
    const int w = (int)(threadIdx.x >> 6);
    for (int k = 0; k < 12; k++) {
        const double t = wave_total_d(g[k]);
        if ((threadIdx.x & 63) == 63) red[w][k] = t;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        double t = 0.0;
        for (int i = 0; i < EX_THREADS / 64; i++) t += red[i][threadIdx.x];
        part[12 * (size_t)blockIdx.x + threadIdx.x] = t;
    }
}

// x[0] + x[stride] + ... + x[(n - 1) stride] in double, added in that order; the loads go
// out 64 at a time (one thread walking the sum would wait out a memory latency per term)
__device__ inline double sum_in_order(const double *x, size_t stride, uint32_t n) {
    constexpr uint32_t B = 64;
    double t = 0.0;
    uint32_t i = 0;
    for (; i + B <= n; i += B) {
        double v[B];
#pragma unroll
        for (uint32_t j = 0; j < B; j++) v[j] = x[stride * (i + j)];
#pragma unroll
        for (uint32_t j = 0; j < B; j++) t += v[j];
    }
    for (; i < n; i++) t += x[stride * i];
    return t;
}

// exposure_adam (cpu:914-923): the sums of the block partials in order, plus the caller's
// extra gradient (then cleared) and the pull towards the identity, one Adam step per value
__global__ void exposure_adam_kernel(const double *part, int nblocks, float *params, float *m1, float *m2,
                                     float *grad, int cam, float step, float bc2, float reg, float b1, float b2,
                                     float eps) {
    const int k = threadIdx.x;
    if (k >= 12) return;
    const double g = sum_in_order(part + k, 12, (uint32_t)nblocks);
    const int i = 12 * cam + k;
    const float ident = (k == 0 || k == 4 || k == 8) ? 1.f : 0.f;
    const float gk = (float)g + (grad ? grad[k] : 0.f) + reg * (params[i] - ident);
    if (grad) grad[k] = 0.f;
    const float m = b1 * m1[i] + (1.0f - b1) * gk;
    const float v = b2 * m2[i] + (1.0f - b2) * gk * gk;
    params[i] -= step * m / (sqrtf(v) / bc2 + eps);
    m1[i] = m;
    m2[i] = v;
}

constexpr double kFix = 281474976710656.0;  // 2^48: fixed-point units of the sky gradient sums

template <int CTRL, int ROW_MASK = 0xf>
__device__ inline unsigned long long dpp_u64(unsigned long long v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, ROW_MASK, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, ROW_MASK, 0xf, false);
    return ((unsigned long long)hi << 32) | lo;
}

// The wave's total of a 64-bit integer, in lane 63 (the DPP steps of a float wave sum:
// partners r^1, r^2, r^7, r^15 within each row of 16, then the row broadcasts); every lane
// of the wave calls it
__device__ inline unsigned long long wave_total_u64(unsigned long long v) {
    v += dpp_u64<0xb1>(v);        // quad_perm [1,0,3,2]
    v += dpp_u64<0x4e>(v);        // quad_perm [2,3,0,1]
    v += dpp_u64<0x141>(v);       // row_half_mirror
    v += dpp_u64<0x140>(v);       // row_mirror
    v += dpp_u64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_u64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// sky_backward (cpu:927-943): T v_img added to the four taps of each pixel's ray. A view
// covers a few thousand texels, so a hundred or more pixels land on each: the wave first
// sums the fixed-point terms of its lanes that share a texel (one pass per distinct texel
// of a tap, led by the lowest pending lane; a DPP reduction leaves the sum in lane 63),
// which adds it with one atomic. Integer sums do not depend on their order, so the
// texture's gradient is the one per-term atomics give.
__global__ __launch_bounds__(256) void sky_grad_kernel(SkyArgs k, uint32_t W, uint32_t H, const float *final_T,
                                                       const float *v_img, unsigned long long *acc) {
    const uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int lane = (int)(threadIdx.x & 63);
    bool live = x < W && y < H;
    uint32_t idx[4] = {0, 0, 0, 0};
    float wt[4] = {0.f, 0.f, 0.f, 0.f}, g[3] = {0.f, 0.f, 0.f};
    if (live) {
        const size_t p = (size_t)y * W + x;
        const float T = final_T[p];
        live = !(T < 1e-4f);
        if (live) {
            g[0] = T * v_img[3 * p];
            g[1] = T * v_img[3 * p + 1];
            g[2] = T * v_img[3 * p + 2];
            sky_taps(k, (float)x, (float)y, idx, wt);
        }
    }
    for (int t = 0; t < 4; t++) {
        bool pend = live && wt[t] != 0.f;
        unsigned long long q[3] = {0ull, 0ull, 0ull};
        if (pend)
            for (int c = 0; c < 3; c++) q[c] = (unsigned long long)__double2ll_rn((double)(wt[t] * g[c]) * kFix);
        for (uint64_t m = __ballot(pend); m; m = __ballot(pend)) {
            const int leader = __builtin_ctzll(m);
            const uint32_t key = (uint32_t)__builtin_amdgcn_readlane((int)idx[t], leader);
            const bool mine = pend && idx[t] == key;
            unsigned long long s[3];
            for (int c = 0; c < 3; c++) s[c] = wave_total_u64(mine ? q[c] : 0ull);
            if (lane == 63)
                for (int c = 0; c < 3; c++) atomicAdd(acc + 3 * (size_t)key + c, s[c]);
            pend = pend && !mine;
        }
    }
}

// The texture's gradient (written out, as the reference leaves it in sky_grad) and its Adam step
__global__ __launch_bounds__(256) void sky_adam_kernel(float *tex, float *m1, float *m2, float *grad,
                                                       const unsigned long long *acc, size_t n, float step, float bc2,
                                                       float b1, float b2, float eps) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float g = (float)((double)(long long)acc[i] * (1.0 / kFix));
    grad[i] = g;
    const float m = b1 * m1[i] + (1.0f - b1) * g;
    const float v = b2 * m2[i] + (1.0f - b2) * g * g;
    tex[i] -= step * m / (sqrtf(v) / bc2 + eps);
    m1[i] = m;
    m2[i] = v;
}

__global__ __launch_bounds__(256) void zero_u64_kernel(unsigned long long *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// scale_ratio_cap (cpu:1295-1309): the largest log-scale is lowered to the middle one plus
// log_max (of equal largest values, the first)
__global__ __launch_bounds__(256) void scale_cap_kernel(float *s, int n, float log_max) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float a = s[3 * i], b = s[3 * i + 1], c = s[3 * i + 2];
    const float mx = fmaxf(a, fmaxf(b, c)), mn = fminf(a, fminf(b, c));
    const float cap = (a + b + c - mx - mn) + log_max;
    if (mx <= cap) return;
    if (a == mx) s[3 * i] = cap;
    else if (b == mx) s[3 * i + 1] = cap;
    else s[3 * i + 2] = cap;
}

// msplat_finalize_depth (cpu:1736-1745): alpha = 1 - T and the expected depth
// depth_num / alpha * inv_scale (0 where alpha <= 1e-4)
__global__ __launch_bounds__(256) void finalize_depth_kernel(const float *d, const float *t, uint32_t n,
                                                             float inv_scale, float *od, float *oa) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float a = 1.f - t[i];
    oa[i] = a;
    od[i] = a > 1e-4f ? d[i] / a * inv_scale : 0.f;
}

}  // namespace

void launch_finalize_depth(const float *depth_num, const float *final_T, uint32_t n, float inv_scale,
                           float *depth_out, float *alpha_out, hipStream_t s) {
    if (n) finalize_depth_kernel<<<(n + 255) / 256, 256, 0, s>>>(depth_num, final_T, n, inv_scale, depth_out, alpha_out);
}

void launch_sky_background(const SkyArgs &k, uint32_t W, uint32_t H, const float bg[3], float *bg_img, hipStream_t s) {
    if (W && H)
        sky_background_kernel<<<dim3((W + 15) / 16, (H + 15) / 16), 256, 0, s>>>(k, W, H, bg[0], bg[1], bg[2], bg_img);
}

void launch_exposure_apply(const float *img, const float *P, size_t n, float *out, hipStream_t s) {
    if (n) exposure_apply_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(img, P, n, out);
}

void launch_prior_loss(const PriorLossArgs &a, float *loss_terms, hipStream_t s) {
    if (!a.width || !a.height) return;
    prior_loss_kernel<<<a.height, 256, 0, s>>>(a);
    if (loss_terms) prior_terms_kernel<<<1, 256, 0, s>>>(a.rows, a.height, 1.0f / (float)(a.width * a.height), loss_terms);
}

size_t exposure_partials(size_t n) { return (n + EX_ITEMS - 1) / EX_ITEMS; }

void launch_exposure_backward_adam(const float *img, const float *v_adj, size_t n, float *v_img, double *part,
                                   float *params, float *m1, float *m2, float *grad, int cam, float step, float bc2,
                                   float reg, float b1, float b2, float eps, hipStream_t s) {
    const size_t nb = exposure_partials(n);
    if (nb) exposure_backward_kernel<<<(unsigned)nb, EX_THREADS, 0, s>>>(img, v_adj, n, params + 12 * (size_t)cam, v_img, part);
    exposure_adam_kernel<<<1, 64, 0, s>>>(part, (int)nb, params, m1, m2, grad, cam, step, bc2, reg, b1, b2, eps);
}

void launch_sky_backward_adam(const SkyArgs &k, uint32_t W, uint32_t H, const float *final_T, const float *v_img,
                              unsigned long long *acc, float *grad, float *tex, float *m1, float *m2, float step,
                              float bc2, float b1, float b2, float eps, hipStream_t s) {
    const size_t n = (size_t)k.w * k.h * 3;
    zero_u64_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(acc, n);
    if (W && H) sky_grad_kernel<<<dim3((W + 15) / 16, (H + 15) / 16), 256, 0, s>>>(k, W, H, final_T, v_img, acc);
    sky_adam_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(tex, m1, m2, grad, acc, n, step, bc2, b1, b2, eps);
}

void launch_scale_cap(float *scales, int n, float log_max, hipStream_t s) {
    if (n > 0) scale_cap_kernel<<<(n + 255) / 256, 256, 0, s>>>(scales, n, log_max);
}

}  // namespace msplat
