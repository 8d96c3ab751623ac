// Exact 3-nearest-neighbour mean distance and Gaussian initialisation
// (msplat_knn3_mean_dist / msplat_init_gaussians, core/cpu/msplat_cpu.cpp:1499-1605).
//
// Points are ordered along a 30-bit Morton curve (LSD radix sort), grouped into boxes of
// KNN_BOX consecutive points and super-boxes of KNN_SUPER boxes, each with its AABB.
// A query seeds its three best squared distances from its Morton neighbours, then visits
// every super-box and box whose AABB can still hold a closer point; culling uses a
// conservative margin, so the three smallest squared distances (computed as
// dx*dx + dy*dy + dz*dz in float, unfused) are exactly those of a brute-force search.
#include "common.h"
#include "kernels.h"

namespace msplat {

constexpr int KNN_BOX = 128;
constexpr int KNN_SUPER = 16;  // boxes per super-box
constexpr int KNN_SEED = 16;   // Morton neighbours per side used to seed the search

__device__ inline uint32_t spread10(uint32_t v) {
    v &= 0x3FF;
    v = (v | (v << 16)) & 0x030000FF;
    v = (v | (v << 8)) & 0x0300F00F;
    v = (v | (v << 4)) & 0x030C30C3;
    v = (v | (v << 2)) & 0x09249249;
    return v;
}

// per-block min/max of the points; one float[6] per block
__global__ __launch_bounds__(256) void bounds_kernel(const float *pts, uint32_t n, float *out) {
    __shared__ float s[6][256];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        for (int a = 0; a < 3; a++) {
            lo[a] = fminf(lo[a], pts[3 * i + a]);
            hi[a] = fmaxf(hi[a], pts[3 * i + a]);
        }
    for (int a = 0; a < 3; a++) {
        s[a][threadIdx.x] = lo[a];
        s[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int a = 0; a < 3; a++) {
                s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + st]);
                s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + st]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// reduces the per-block bounds to one (written to out[0..6))
__global__ __launch_bounds__(256) void bounds_final_kernel(const float *blk, int nblk, float *out) {
    __shared__ float s[6][256];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nblk; b += 256)
        for (int a = 0; a < 3; a++) {
            lo[a] = fminf(lo[a], blk[6 * b + a]);
            hi[a] = fmaxf(hi[a], blk[6 * b + 3 + a]);
        }
    for (int a = 0; a < 3; a++) {
        s[a][threadIdx.x] = lo[a];
        s[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int a = 0; a < 3; a++) {
                s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + st]);
                s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + st]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void morton_kernel(const float *pts, uint32_t n, const float *bb, uint64_t *codes,
                                                    int32_t *idx) {
    const float lo[3] = {bb[0], bb[1], bb[2]}, hi[3] = {bb[3], bb[4], bb[5]};
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        const float ext = hi[a] - lo[a];
        float t = ext > 0.f ? (pts[3 * i + a] - lo[a]) / ext : 0.f;
        q[a] = (uint32_t)fminf(fmaxf(t * 1023.f, 0.f), 1023.f);
    }
    codes[i] = (uint64_t)(spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2));
    idx[i] = (int32_t)i;
}

__global__ __launch_bounds__(256) void gather_points_kernel(const float *pts, const int32_t *order, uint32_t n,
                                                           float4 *sorted) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int32_t j = order[i];
    sorted[i] = make_float4(pts[3 * j], pts[3 * j + 1], pts[3 * j + 2], 0.f);
}

// AABB of box b (KNN_BOX consecutive sorted points); lo in [6b..], hi in [6b+3..]
__global__ __launch_bounds__(256) void box_bounds_kernel(const float4 *sp, uint32_t n, int nbox, float *box) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= nbox) return;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const uint32_t e = min(n, (uint32_t)(b + 1) * KNN_BOX);
    for (uint32_t i = (uint32_t)b * KNN_BOX; i < e; i++) {
        const float4 p = sp[i];
        lo[0] = fminf(lo[0], p.x); lo[1] = fminf(lo[1], p.y); lo[2] = fminf(lo[2], p.z);
        hi[0] = fmaxf(hi[0], p.x); hi[1] = fmaxf(hi[1], p.y); hi[2] = fmaxf(hi[2], p.z);
    }
    for (int a = 0; a < 3; a++) {
        box[6 * b + a] = lo[a];
        box[6 * b + 3 + a] = hi[a];
    }
}

__global__ __launch_bounds__(256) void super_bounds_kernel(const float *box, int nbox, int nsup, float *sup) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= nsup) return;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = s * KNN_SUPER; b < min(nbox, (s + 1) * KNN_SUPER); b++)
        for (int a = 0; a < 3; a++) {
            lo[a] = fminf(lo[a], box[6 * b + a]);
            hi[a] = fmaxf(hi[a], box[6 * b + 3 + a]);
        }
    for (int a = 0; a < 3; a++) {
        sup[6 * s + a] = lo[a];
        sup[6 * s + 3 + a] = hi[a];
    }
}

__device__ inline void insert3(float d, float &b0, float &b1, float &b2) {
    if (d >= b2) return;
    if (d < b1) {
        b2 = b1;
        if (d < b0) {
            b1 = b0;
            b0 = d;
        } else {
            b1 = d;
        }
    } else {
        b2 = d;
    }
}

// squared distance from q to an AABB, in double, shrunk by a relative margin so that a
// float-rounded point distance can never be smaller than the culling bound
__device__ inline bool box_may_hold(const float *bx, float qx, float qy, float qz, float best) {
    const double dx = fmax(0.0, fmax((double)bx[0] - qx, (double)qx - bx[3]));
    const double dy = fmax(0.0, fmax((double)bx[1] - qy, (double)qy - bx[4]));
    const double dz = fmax(0.0, fmax((double)bx[2] - qz, (double)qz - bx[5]));
    const double d2 = dx * dx + dy * dy + dz * dz;
    return d2 * (1.0 - 1e-5) <= (double)best;
}

__global__ __launch_bounds__(256) void knn3_kernel(const float4 *sp, const int32_t *order, uint32_t n,
                                                  const float *box, int nbox, const float *sup, int nsup,
                                                  float *mean_dist) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 q = sp[i];
    float b0 = INFINITY, b1 = INFINITY, b2 = INFINITY;
    auto visit = [&](uint32_t j) {
        if (j == i) return;
        const float4 p = sp[j];
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        insert3(dx * dx + dy * dy + dz * dz, b0, b1, b2);
    };
    // seed from the Morton neighbours, then the culled exhaustive search; the seed range
    // is skipped there so that no point is inserted twice
    const uint32_t s0 = i > KNN_SEED ? i - KNN_SEED : 0, s1 = min(n, i + KNN_SEED + 1);
    for (uint32_t j = s0; j < s1; j++) visit(j);
    for (int s = 0; s < nsup; s++) {
        if (!box_may_hold(sup + 6 * s, q.x, q.y, q.z, b2)) continue;
        for (int b = s * KNN_SUPER; b < min(nbox, (s + 1) * KNN_SUPER); b++) {
            if (!box_may_hold(box + 6 * b, q.x, q.y, q.z, b2)) continue;
            const uint32_t e = min(n, (uint32_t)(b + 1) * KNN_BOX);
            for (uint32_t j = (uint32_t)b * KNN_BOX; j < e; j++)
                if (j < s0 || j >= s1) visit(j);
        }
    }
    float sum = 0.f, cnt = 0.f;
    if (isfinite(b0)) { sum += sqrtf(b0); cnt += 1.f; }
    if (isfinite(b1)) { sum += sqrtf(b1); cnt += 1.f; }
    if (isfinite(b2)) { sum += sqrtf(b2); cnt += 1.f; }
    mean_dist[order[i]] = cnt > 0.f ? sum / cnt : 1.f;
}

// log-scales, Shoemake rotations from the PCG stream, DC colour, constant opacity
__global__ __launch_bounds__(256) void init_gaussians_kernel(const float *md, const uint8_t *rgb, uint32_t n,
                                                            uint32_t seed, float logit, float *scales, float *quats,
                                                            float *dc, float *opac) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float ls = log_cr(fmaxf(md[i], 1e-7f));
    scales[3 * i] = ls;
    scales[3 * i + 1] = ls;
    scales[3 * i + 2] = ls;
    const float u = rand_uniform(seed, 3u * i), v = rand_uniform(seed, 3u * i + 1u), w = rand_uniform(seed, 3u * i + 2u);
    const float tau = 6.28318530717958647692f;
    quats[4 * i + 0] = sqrtf(1.f - u) * sin_cr(tau * v);
    quats[4 * i + 1] = sqrtf(1.f - u) * cos_cr(tau * v);
    quats[4 * i + 2] = sqrtf(u) * sin_cr(tau * w);
    quats[4 * i + 3] = sqrtf(u) * cos_cr(tau * w);
    for (int k = 0; k < 3; k++) dc[3 * i + k] = ((float)rgb[3 * i + k] * (1.f / 255.f) - 0.5f) / SH_C0;
    opac[i] = logit;
}

int knn3_launch(const float *pts, uint32_t n, float *mean_dist, KnnScratch &w, hipStream_t s) {
    if (n == 0) return 0;
    const int nb = (int)((n + 255) / 256);
    const int nbb = nb < 1024 ? nb : 1024;
    bounds_kernel<<<nbb, 256, 0, s>>>(pts, n, w.bounds);
    bounds_final_kernel<<<1, 256, 0, s>>>(w.bounds, nbb, w.bounds + 6 * 1024);
    morton_kernel<<<nb, 256, 0, s>>>(pts, n, w.bounds + 6 * 1024, w.codes, w.order);
    fill_u32<<<1, 1, 0, s>>>(w.n_dev, n);
    for (int shift = 0; shift < 30; shift += 8)
        launch_radix_pass(w.codes, w.order, w.codes_tmp, w.order_tmp, w.n_dev, n, shift, w.hist, w.digit_total, s),
            std::swap(w.codes, w.codes_tmp), std::swap(w.order, w.order_tmp);
    const int nbox = (int)((n + KNN_BOX - 1) / KNN_BOX), nsup = (nbox + KNN_SUPER - 1) / KNN_SUPER;
    gather_points_kernel<<<nb, 256, 0, s>>>(pts, w.order, n, w.sorted);
    box_bounds_kernel<<<(nbox + 255) / 256, 256, 0, s>>>(w.sorted, n, nbox, w.box);
    super_bounds_kernel<<<(nsup + 255) / 256, 256, 0, s>>>(w.box, nbox, nsup, w.sup);
    knn3_kernel<<<nb, 256, 0, s>>>(w.sorted, w.order, n, w.box, nbox, w.sup, nsup, mean_dist);
    return 0;
}

size_t knn_scratch_bytes(uint32_t n) {
    const size_t nbox = (n + KNN_BOX - 1) / KNN_BOX, nsup = (nbox + KNN_SUPER - 1) / KNN_SUPER;
    return 6 * 1025 * 4 + 2 * 8 * (size_t)n + 2 * 4 * (size_t)n + 16 * (size_t)n + 6 * 4 * (nbox + nsup) +
           4 * radix_hist_words(n) + 4 * 256 + 64 + 16 * 8;
}

void knn_scratch_carve(void *base, uint32_t n, KnnScratch &w) {
    const size_t nbox = (n + KNN_BOX - 1) / KNN_BOX, nsup = (nbox + KNN_SUPER - 1) / KNN_SUPER;
    char *p = (char *)base;
    auto take = [&](size_t bytes) {
        char *r = p;
        p += (bytes + 15) & ~(size_t)15;
        return (void *)r;
    };
    w.bounds = (float *)take(6 * 1025 * 4);
    w.codes = (uint64_t *)take(8 * (size_t)n);
    w.codes_tmp = (uint64_t *)take(8 * (size_t)n);
    w.order = (int32_t *)take(4 * (size_t)n);
    w.order_tmp = (int32_t *)take(4 * (size_t)n);
    w.sorted = (float4 *)take(16 * (size_t)n);
    w.box = (float *)take(6 * 4 * nbox);
    w.sup = (float *)take(6 * 4 * nsup);
    w.hist = (uint32_t *)take(4 * radix_hist_words(n));
    w.digit_total = (uint32_t *)take(4 * 256);
    w.n_dev = (uint32_t *)take(64);
}

void launch_init_gaussians(const float *md, const uint8_t *rgb, uint32_t n, uint32_t seed, float logit, float *scales,
                           float *quats, float *dc, float *opac, hipStream_t s) {
    if (n == 0) return;
    init_gaussians_kernel<<<(n + 255) / 256, 256, 0, s>>>(md, rgb, n, seed, logit, scales, quats, dc, opac);
}

}  // namespace msplat
