// The ground-truth image pipeline before the training step: area resampling (the
// pyramid levels of a training image) and Brown-Conrady undistortion, one thread per
// output pixel. Both keep the reference's operation order (msplat_resize_area and
// msplat_undistort, core/cpu/msplat_cpu.cpp:1609-1669), so results are bit for bit the
// CPU backend's: the whole library is built without floating-point contraction and
// divisions are correctly rounded.
#include "context.h"

using namespace msplat;

namespace {

// msplat_resize_area (cpu:1609-1639): each output pixel averages the source pixels its
// footprint [gx * sw/dw, (gx+1) * sw/dw) x [gy * sh/dh, (gy+1) * sh/dh) overlaps, weighted
// by the overlap area, rows outer and columns inner. A u8 source is RGBA (4 bytes per
// pixel, alpha ignored) scaled by 1/255; a float source is RGB.
__global__ __launch_bounds__(256) void resize_area_kernel(const uint8_t *u8, const float *fsrc, int sw, int sh,
                                                          float *out, int dw, int dh, float scale_x,
                                                          float scale_y) {
    const int gx = blockIdx.x * 16 + (threadIdx.x & 15), gy = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (gx >= dw || gy >= dh) return;
    const float sy0 = (float)gy * scale_y, sy1 = (float)(gy + 1) * scale_y;
    const int iy0 = (int)sy0, iy1 = min((int)ceilf(sy1), sh);
    const float sx0 = (float)gx * scale_x, sx1 = (float)(gx + 1) * scale_x;
    const int ix0 = (int)sx0, ix1 = min((int)ceilf(sx1), sw);
    float r = 0.f, g = 0.f, b = 0.f, area = 0.f;
    for (int iy = iy0; iy < iy1; iy++) {
        const float wy = fminf((float)(iy + 1), sy1) - fmaxf((float)iy, sy0);
        for (int ix = ix0; ix < ix1; ix++) {
            const float wx = fminf((float)(ix + 1), sx1) - fmaxf((float)ix, sx0);
            const size_t s = (size_t)iy * sw + ix;
            float cr, cg, cb;
            if (u8) {
                const uchar4 p = *(const uchar4 *)(u8 + 4 * s);
                cr = (float)p.x * (1.f / 255.f);
                cg = (float)p.y * (1.f / 255.f);
                cb = (float)p.z * (1.f / 255.f);
            } else {
                cr = fsrc[3 * s];
                cg = fsrc[3 * s + 1];
                cb = fsrc[3 * s + 2];
            }
            const float w = wx * wy;
            r = r + w * cr;
            g = g + w * cg;
            b = b + w * cb;
            area = area + w;
        }
    }
    const size_t o = 3 * ((size_t)gy * dw + gx);
    out[o] = r / area;
    out[o + 1] = g / area;
    out[o + 2] = b / area;
}

// msplat_undistort (cpu:1641-1669): the undistorted crop's pixel centre (gx + roi_x + 0.5,
// gy + roi_y + 0.5) goes through the Brown-Conrady model (k1, k2, k3 radial, p1, p2
// tangential) to a source position, sampled bilinearly with clamp-to-edge.
struct UndistortArgs {
    const float *in;
    float *out;
    int sw, sh, dw, dh, roi_x, roi_y;
    float fx, fy, cx, cy, k1, k2, p1, p2, k3;
};

__global__ __launch_bounds__(256) void undistort_kernel(UndistortArgs a) {
    const int gx = blockIdx.x * 16 + (threadIdx.x & 15), gy = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (gx >= a.dw || gy >= a.dh) return;
    const float x = ((float)(gx + a.roi_x) + 0.5f - a.cx) / a.fx;
    const float y = ((float)(gy + a.roi_y) + 0.5f - a.cy) / a.fy;
    const float r2 = x * x + y * y;
    const float radial = 1.f + a.k1 * r2 + a.k2 * r2 * r2 + a.k3 * r2 * r2 * r2;
    const float xd = x * radial + 2.f * a.p1 * x * y + a.p2 * (r2 + 2.f * x * x);
    const float yd = y * radial + a.p1 * (r2 + 2.f * y * y) + 2.f * a.p2 * x * y;
    const float sx = xd * a.fx + a.cx - 0.5f, sy = yd * a.fy + a.cy - 0.5f;
    const float fsx = floorf(sx), fsy = floorf(sy);
    const int x0 = (int)fsx, y0 = (int)fsy;
    const float ax = sx - fsx, ay = sy - fsy;
    const int xa = min(max(x0, 0), a.sw - 1), xb = min(max(x0 + 1, 0), a.sw - 1);
    const int ya = min(max(y0, 0), a.sh - 1), yb = min(max(y0 + 1, 0), a.sh - 1);
    const float *c00 = a.in + 3 * ((size_t)ya * a.sw + xa), *c10 = a.in + 3 * ((size_t)ya * a.sw + xb);
    const float *c01 = a.in + 3 * ((size_t)yb * a.sw + xa), *c11 = a.in + 3 * ((size_t)yb * a.sw + xb);
    const size_t o = 3 * ((size_t)gy * a.dw + gx);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float top = c00[c] * (1.f - ax) + c10[c] * ax;
        const float bottom = c01[c] * (1.f - ax) + c11[c] * ax;
        a.out[o + c] = top * (1.f - ay) + bottom * ay;
    }
}

}  // namespace

extern "C" {

int msplat_hip_resize_area(msplat_hip_ctx *ctx, const void *src, int src_is_u8, int sw, int sh, float *dst, int dw,
                           int dh) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_replays(ctx));
    if (dw <= 0 || dh <= 0) return MSPLAT_OK;  // as the reference: nothing to write
    if (!src || !dst) return set_error(MSPLAT_ERR_ARG, "src or dst is NULL");
    if (sw <= 0 || sh <= 0) return set_error(MSPLAT_ERR_ARG, "source size must be positive");
    const float scale_x = (float)sw / (float)dw, scale_y = (float)sh / (float)dh;
    const dim3 grid((dw + 15) / 16, (dh + 15) / 16);
    resize_area_kernel<<<grid, 256, 0, ctx->stream>>>(src_is_u8 ? (const uint8_t *)src : nullptr,
                                                      src_is_u8 ? nullptr : (const float *)src, sw, sh, dst, dw, dh,
                                                      scale_x, scale_y);
    HIPCHECK(hipGetLastError());
    return MSPLAT_OK;
}

int msplat_hip_undistort(msplat_hip_ctx *ctx, const float *src, int sw, int sh, const float intr[4],
                         const float dist[5], int roi_x, int roi_y, float *dst, int dw, int dh) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    CHECKED(settle_replays(ctx));
    if (dw <= 0 || dh <= 0) return MSPLAT_OK;
    if (!src || !dst || !intr || !dist) return set_error(MSPLAT_ERR_ARG, "NULL argument");
    if (sw <= 0 || sh <= 0) return set_error(MSPLAT_ERR_ARG, "source size must be positive");
    UndistortArgs a;
    a.in = src;
    a.out = dst;
    a.sw = sw;
    a.sh = sh;
    a.dw = dw;
    a.dh = dh;
    a.roi_x = roi_x;
    a.roi_y = roi_y;
    a.fx = intr[0];
    a.fy = intr[1];
    a.cx = intr[2];
    a.cy = intr[3];
    a.k1 = dist[0];
    a.k2 = dist[1];
    a.p1 = dist[2];
    a.p2 = dist[3];
    a.k3 = dist[4];
    const dim3 grid((dw + 15) / 16, (dh + 15) / 16);
    undistort_kernel<<<grid, 256, 0, ctx->stream>>>(a);
    HIPCHECK(hipGetLastError());
    return MSPLAT_OK;
}

}  // extern "C"
