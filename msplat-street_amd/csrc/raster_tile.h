// Per-tile front-to-back alpha compositing (forward) and its backward, as device
// functions shared by the one-view kernels (one workgroup per tile) and the persistent
// batched kernels (workgroups pull (view, tile) items from a queue).
//
// Forward: 256 threads per 16x16 tile; each wave64 owns an 8x8 quadrant so that a
// wave's pixels are spatially tight. Gaussian records of the tile's sorted list are
// staged through LDS in batches (the LDS batch depth is a kernel variant).
//
// Backward without global atomics: per gaussian of a batch, the wave's contributions
// are summed (DPP within rows of 16 lanes, permlane swaps across rows) only when some
// lane contributed (ballot), written to a per-wave LDS partial, and the partials are
// summed in wave order into the intersection's gradient slot. Slots are laid out in the pre-sort order of the intersections
// (record-major, tiles row-major), so the per-gaussian reduction reads each record's
// slots contiguously and in the reference's tile order.
#pragma once

#include "common.h"
#include "kernels.h"

namespace msplat {

// The pixels of one view: `image`, `final_T`, `final_idx` and `v_image` point at the
// view's first pixel.
struct TileView {
    uint32_t width, height;
    int tiles_x;
    float bg[3];
    float *image, *final_T;
    int32_t *final_idx;
    const float *v_image;
    // prior-guided (aux) frames: the per-pixel background, the expected-depth sum, and
    // the prior losses' gradients dL/dD, dL/dA (2 per pixel)
    const float *bg_img = nullptr;
    float *out_depth = nullptr;
    const float *v_aux = nullptr;
};

// Buffers shared by every view: ranges over global tile ids, the sorted record ids,
// per-record splats and tile rectangles (in the record's view), and gradient slots.
struct TileData {
    const int32_t *tile_start, *gid;
    uint32_t gid_mask;  // record id = gid[k] & gid_mask
    const float4 *splat_a, *splat_b;
    const float *splat_blue;
    const uint2 *rect;
    const uint32_t *offsets;
    float4 *slot_a, *slot_b;
    float *slot_c;
    uint8_t *slot_f;  // 1: the slot's nine values were written; 0 (set before the launch): all zero
    unsigned long long *slots_written;  // null, or a running count of written slots
    bool cull;  // skip gaussians that splat_misses_tile proves invisible in the tile
    // aux frames: per record view depth, and the intersections' depth gradient slots
    const float *depths = nullptr;
    float *slot_d = nullptr;
};

// True when the splat composites no pixel of the tile whose top-left pixel is (x0, y0):
// the tile lies outside the bounding box of the ellipse {0.5 d^T Q d <= L'}, with Q the
// conic and L' above the largest sigma that passes both sigma < SIGMA_MAX and
// opacity exp(-sigma) >= ALPHA_MIN. The pixel loops compute -log2(e) sigma from the
// scaled conic (one more rounding per coefficient) to a relative error below 24 u kappa
// (u = 2^-24, kappa the conic's condition number) and alpha to a few ulp, so the test is
// made only for kappa < 1024 (24 u kappa < 1.5e-3) and L' is the exact level raised by
// 2^-8 relative plus 2^-8 absolute; the box is widened by 0.1% plus 0.01 px for the
// rounding of its own arithmetic. Any pair it rejects would fail the validity test in
// every pixel, and an invalid pair changes nothing (it adds exact zeros), so skipping
// it leaves every output bit unchanged.
__device__ inline bool splat_misses_tile(const float4 &A, const float4 &B, float x0, float y0) {
    const float a = A.z, b = A.w, c = B.x;
    const float det = a * c - b * b, tr = a + c;
    if (!(det > 0.f) || !(tr * tr < 1024.f * det)) return false;  // (a+c)^2/det = kappa + 2 + 1/kappa
    const float L = fminf(SIGMA_MAX, __logf(B.y * (1.f / ALPHA_MIN)));
    if (!(L > 1.f / 64.f)) return false;
    const float k2 = 2.f * (L * (1.f + 1.f / 256.f) + 1.f / 256.f) / det;
    const float hx = sqrtf(k2 * c) * 1.001f + 0.01f, hy = sqrtf(k2 * a) * 1.001f + 0.01f;
    const float x1 = x0 + (float)(TILE - 1), y1 = y0 + (float)(TILE - 1);
    return A.x + hx < x0 || A.x - hx > x1 || A.y + hy < y0 || A.y - hy > y1;
}

// The pixel loops work with p = -log2(e) sigma, the exponent of alpha's power of two
// (alpha = opacity 2^p, one v_exp_f32), from a conic scaled when it is staged in LDS:
// a' = -log2(e)/2 a, b' = -log2(e) b, c' = -log2(e)/2 c, so that
// p = (a' dx) dx + (c' dy) dy + (b' dx) dy; sigma >= 0 is p <= 0 and sigma < SIGMA_MAX
// is p > P_MIN. One fixed evaluation order is shared by the forward and backward loops,
// so that both make the same validity decisions. adx2 = (a' dx) dx and bdx = b' dx are
// per record, which the backward hoists out of its pixel rows.
constexpr float POW_HALF = -0.72134752044448170f;  // -log2(e) / 2
constexpr float POW_ONE = -1.44269504088896341f;   // -log2(e)
__device__ constexpr float P_MIN = SIGMA_MAX * POW_ONE;
__device__ inline float4 staged_a(const float4 &A) { return make_float4(A.x, A.y, A.z * POW_HALF, A.w * POW_ONE); }
__device__ inline float splat_power(float adx2, float bdx, float c, float dy) {
    return __builtin_fmaf(bdx, dy, __builtin_fmaf(c * dy, dy, adx2));
}
__device__ inline float splat_alpha(float opacity, float p) {
    return fminf(ALPHA_MAX, opacity * __builtin_amdgcn_exp2f(p));
}
__device__ inline bool splat_valid(float p, float alpha) { return p <= 0.f && p > P_MIN && alpha >= ALPHA_MIN; }

// Position of the calling thread's item among the workgroup's kept items (in thread
// order) and, in `total`, their number. Every thread of the NW-wave workgroup calls it;
// `wcount` must not be written again before every wave has returned (callers alternate
// two buffers or pass a barrier in between).
template <int NW>
__device__ inline int kept_position(bool keep, int *wcount, int &total) {
    const uint64_t bal = __ballot(keep);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wcount[w] = __popcll(bal);
    __syncthreads();
    int off = 0;
    total = 0;
#pragma unroll
    for (int i = 0; i < NW; i++) {
        const int cnt = wcount[i];
        off += i < w ? cnt : 0;
        total += cnt;
    }
    return off + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
}

// nd_rasterize_forward (core/cpu/msplat_cpu.cpp:574-626) for global tile `gt`, which is
// tile `lt` of view `v`. Every thread of the 256-thread workgroup calls it; FB gaussian
// records are staged in LDS per batch (FB <= 256: one record per thread).
// With AUX (rasterize_forward_aux, cpu:574-626 with f.aux): also the expected-depth sum
// sum_i z_i alpha_i T_i into out_depth, over the per-pixel background bg_img.
template <int FB, bool AUX = false>
__device__ inline void forward_tile(const TileView &v, const TileData &d, int gt, int lt) {
#pragma clang fp contract(fast)
    static_assert(FB <= 256, "one staged record per thread");
    __shared__ float4 la[FB], lb[FB];
    __shared__ float lc[FB];
    __shared__ int lk[FB];  // the staged record's position in the tile's list
    __shared__ float lz[AUX ? FB : 1];  // the record's view depth
    __shared__ int wcount[4];
    const int tx = lt % v.tiles_x, ty = lt / v.tiles_x;
    const float x0 = (float)(tx * TILE), y0 = (float)(ty * TILE);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int px = tx * TILE + (w & 1) * 8 + (lane & 7);
    const int py = ty * TILE + (w >> 1) * 8 + (lane >> 3);
    const bool inside = px < (int)v.width && py < (int)v.height;
    const int s = d.tile_start[gt], e = d.tile_start[gt + 1];
    const float fx = (float)px, fy = (float)py;
    float T = 1.f, cr = 0.f, cg = 0.f, cb = 0.f, depth = 0.f;
    int last = s - 1;
    bool done = !inside;
    for (int base = s; base < e; base += FB) {
        if (__syncthreads_count(done) == 256) break;
        // stage the batch's records that can reach the tile, in list order
        const int k = base + threadIdx.x;
        bool keep = false;
        float4 A0, B0;
        float C0 = 0.f, Z0 = 0.f;
        if (threadIdx.x < FB && k < e) {
            const int g = (int)((uint32_t)d.gid[k] & d.gid_mask);
            A0 = d.splat_a[g];
            B0 = d.splat_b[g];
            C0 = d.splat_blue[g];
            if (AUX) Z0 = d.depths[g];
            keep = !(d.cull && splat_misses_tile(A0, B0, x0, y0));
        }
        int n;
        const int pos = kept_position<4>(keep, wcount, n);
        if (keep) {  // colours stored clamped: max(raw + 0.5, 0), as composited
            la[pos] = staged_a(A0);
            lb[pos] = make_float4(B0.x * POW_HALF, B0.y, fmaxf(B0.z + 0.5f, 0.f), fmaxf(B0.w + 0.5f, 0.f));
            lc[pos] = fmaxf(C0 + 0.5f, 0.f);
            lk[pos] = k;
            if (AUX) lz[pos] = Z0;
        }
        // pad to a multiple of 8 with records of zero opacity (alpha 0 < ALPHA_MIN: they
        // composite nothing), so that the loop below runs whole groups of 8
        const int n8 = (n + 7) & ~7;
        if ((int)threadIdx.x >= n && (int)threadIdx.x < n8) {
            la[threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
            lb[threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
            lc[threadIdx.x] = 0.f;
            lk[threadIdx.x] = 0;
            if (AUX) lz[threadIdx.x] = 0.f;
        }
        __syncthreads();
        if (__ballot(!done) != 0) {
            for (int j0 = 0; j0 < n8; j0 += 8) {
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    // predicated compositing step: skip (out of range / too faint), accept,
                    // or terminate (transmittance would drop to T_MIN: last = k - 1, as the
                    // CPU)
                    const int j = j0 + u;
                    const float4 A = la[j], B = lb[j];
                    const float dx = A.x - fx, dy = A.y - fy;
                    const float pw = splat_power((A.z * dx) * dx, A.w * dx, B.x, dy);
                    const float alpha = splat_alpha(B.y, pw);
                    const bool valid = !done && splat_valid(pw, alpha);
                    const float nT = T * (1.f - alpha);
                    const bool term = valid && nT <= T_MIN;
                    const bool accept = valid && !term;
                    const float vis = accept ? alpha * T : 0.f;
                    cr += B.z * vis;
                    cg += B.w * vis;
                    cb += lc[j] * vis;
                    if (AUX) depth += lz[j] * vis;
                    T = accept ? nT : T;
                    const int kj = lk[j];
                    last = term ? kj - 1 : (accept ? kj : last);
                    done = done || term;
                }
                if (__ballot(!done) == 0) break;
            }
        }
    }
    if (inside) {
        const size_t pid = (size_t)py * v.width + px;
        v.final_T[pid] = T;
        v.final_idx[pid] = last;
        float bg0 = v.bg[0], bg1 = v.bg[1], bg2 = v.bg[2];
        if (AUX) {
            bg0 = v.bg_img[3 * pid];
            bg1 = v.bg_img[3 * pid + 1];
            bg2 = v.bg_img[3 * pid + 2];
            v.out_depth[pid] = depth;
        }
        v.image[3 * pid] = fminf(fmaxf(bg0 * T + cr, 0.f), 1.f);
        v.image[3 * pid + 1] = fminf(fmaxf(bg1 * T + cg, 0.f), 1.f);
        v.image[3 * pid + 2] = fminf(fmaxf(bg2 * T + cb, 0.f), 1.f);
    }
}

// ── Depth-chunked compositing for views with few tiles ─────────────────────────
// A view with few tiles runs few workgroups, each over a long list. The chunked form
// splits every tile's list into K consecutive parts of equal length (chunk c of a list
// [s, e) is [s + c len / K, s + (c + 1) len / K)) and runs one workgroup per (tile,
// chunk), in four launches:
//  1. each chunk's per-pixel product of (1 - alpha) over its valid records, stopped once
//     it reaches T_MIN (every later chunk then starts at or below T_MIN);
//  2. per pixel, the exclusive prefix products in chunk order: the transmittance at each
//     chunk's start;
//  3. each chunk composites front to back from that transmittance with the reference's
//     absolute termination test T (1 - alpha) <= T_MIN; a pixel whose starting
//     transmittance is already at or below T_MIN has terminated in an earlier chunk and
//     composites nothing; a chunk records its colour sum, its end transmittance (negated
//     when the pixel terminated in it) and its last index;
//  4. a merge per pixel: the chunks' colours summed in order up to the terminating
//     chunk, the final transmittance and last contributor.
// The result is the sequential composite of rasterize_forward (cpu:574-626) up to float
// rounding of the transmittance products and colour sums. The backward first computes,
// per pixel and chunk, S at the chunk's end (the background term plus the colours of the
// later chunks up to the terminating one), then runs each chunk back to front from its
// end transmittance and that S.
struct ChunkView {
    int k, K;           // this chunk, chunks per tile
    size_t npix;        // pixels of the view (the stride between chunk planes)
    float *prod;        // [K][npix] the chunk's product of (1 - alpha); after the prefix pass the
                        // transmittance at the chunk's start; before a backward, S at its end
    float4 *state;      // [K][npix] (colour r, g, b, end transmittance; negated: terminated here)
    int32_t *last;      // [K][npix] last index the chunk composited (or k - 1 at termination), -1: none
};

__device__ inline void chunk_range(int s, int e, int k, int K, int &cs, int &ce) {
    const long len = e - s;
    cs = s + (int)(len * k / K);
    ce = s + (int)(len * (k + 1) / K);
}

// Stages records [base, base + FB) of [s, e) that can reach the tile into LDS, in list
// order, padded with zero-opacity records to a multiple of 8; returns the padded count.
template <int FB>
__device__ inline int stage_forward_batch(const TileData &d, int base, int e, float x0, float y0, float4 *la,
                                          float4 *lb, float *lc, int *lk, int *wcount) {
    const int k = base + threadIdx.x;
    bool keep = false;
    float4 A0, B0;
    float C0 = 0.f;
    if (threadIdx.x < FB && k < e) {
        const int g = (int)((uint32_t)d.gid[k] & d.gid_mask);
        A0 = d.splat_a[g];
        B0 = d.splat_b[g];
        C0 = d.splat_blue[g];
        keep = !(d.cull && splat_misses_tile(A0, B0, x0, y0));
    }
    int n;
    const int pos = kept_position<4>(keep, wcount, n);
    if (keep) {
        la[pos] = staged_a(A0);
        lb[pos] = make_float4(B0.x * POW_HALF, B0.y, fmaxf(B0.z + 0.5f, 0.f), fmaxf(B0.w + 0.5f, 0.f));
        lc[pos] = fmaxf(C0 + 0.5f, 0.f);
        lk[pos] = k;
    }
    const int n8 = (n + 7) & ~7;
    if ((int)threadIdx.x >= n && (int)threadIdx.x < n8) {
        la[threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
        lb[threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
        lc[threadIdx.x] = 0.f;
        lk[threadIdx.x] = 0;
    }
    __syncthreads();
    return n8;
}

// Launch 1 (PRODUCT) or 2 of the chunked forward for chunk ch.k of global tile gt.
template <int FB, bool PRODUCT>
__device__ inline void forward_chunk(const TileView &v, const TileData &d, const ChunkView &ch, int gt, int lt) {
#pragma clang fp contract(fast)
    static_assert(FB <= 256, "one staged record per thread");
    __shared__ float4 la[FB], lb[FB];
    __shared__ float lc[FB];
    __shared__ int lk[FB];
    __shared__ int wcount[4];
    const int tx = lt % v.tiles_x, ty = lt / v.tiles_x;
    const float x0 = (float)(tx * TILE), y0 = (float)(ty * TILE);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int px = tx * TILE + (w & 1) * 8 + (lane & 7);
    const int py = ty * TILE + (w >> 1) * 8 + (lane >> 3);
    const bool inside = px < (int)v.width && py < (int)v.height;
    const size_t pid = inside ? (size_t)py * v.width + px : 0;
    int s, e;
    chunk_range(d.tile_start[gt], d.tile_start[gt + 1], ch.k, ch.K, s, e);
    const float fx = (float)px, fy = (float)py;
    float T = 1.f, cr = 0.f, cg = 0.f, cb = 0.f;
    if (!PRODUCT && inside) T = ch.prod[(size_t)ch.k * ch.npix + pid];  // at the chunk's start
    int last = -1;
    bool done = !inside || (!PRODUCT && !(T > T_MIN)), term = false;
    for (int base = s; base < e; base += FB) {
        if (__syncthreads_count(done) == 256) break;
        const int n8 = stage_forward_batch<FB>(d, base, e, x0, y0, la, lb, lc, lk, wcount);
        if (__ballot(!done) == 0) continue;
        for (int j0 = 0; j0 < n8; j0 += 8) {
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int j = j0 + u;
                const float4 A = la[j], B = lb[j];
                const float dx = A.x - fx, dy = A.y - fy;
                const float pw = splat_power((A.z * dx) * dx, A.w * dx, B.x, dy);
                const float alpha = splat_alpha(B.y, pw);
                const bool valid = !done && splat_valid(pw, alpha);
                const float nT = T * (1.f - alpha);
                if (PRODUCT) {  // once T <= T_MIN every later chunk starts terminated: stop there
                    T = valid ? nT : T;
                    done = done || !(T > T_MIN);
                } else {
                    const bool t = valid && nT <= T_MIN;
                    const bool accept = valid && !t;
                    const float vis = accept ? alpha * T : 0.f;
                    cr += B.z * vis;
                    cg += B.w * vis;
                    cb += lc[j] * vis;
                    T = accept ? nT : T;
                    const int kj = lk[j];
                    last = t ? kj - 1 : (accept ? kj : last);
                    term = term || t;
                    done = done || t;
                }
            }
            if (__ballot(!done) == 0) break;
        }
    }
    if (!inside) return;
    const size_t o = (size_t)ch.k * ch.npix + pid;
    if (PRODUCT) {
        ch.prod[o] = T;
    } else {
        ch.state[o] = make_float4(cr, cg, cb, term ? -T : T);
        ch.last[o] = last;
    }
}

// rasterize_backward (core/cpu/msplat_cpu.cpp:959-1031), slot values per intersection:
// a = (v_x, v_y, v_conic.x, v_conic.y), b = (v_conic.z, v_r, v_g, v_b), c = v_logit.
//
// PPT pixels per thread: 4/PPT waves per tile, wave w owning rows [4 PPT w, 4 PPT (w+1))
// and lane l the pixels (l % 16, 4 PPT w + l / 16 + 4 k), k < PPT.
//
// Per pixel the back-to-front recursion keeps T and one running scalar
//   S = <buffer + T_final bg, v_out>,  so  v_alpha = T <rgb, v_out> - S / (1 - alpha),
// and the per-gaussian gradients are assembled after the reduction from sufficient
// statistics of v_sigma = -alpha v_alpha:
//   v_conic = 1/2 (sum vs dx^2, sum vs dx dy, sum vs dy^2),
//   v_xy = (a sum vs dx + b sum vs dy, b sum vs dx + c sum vs dy),
//   v_logit = -(1 - o) sum vs, v_rgb = sum fac v_out (channels with raw + 0.5 >= 0).
// A lane first adds its PPT pixels; the nine values are then summed over the wave
// (wave_sums9) and lanes 0..8 write the wave's totals to LDS, to be added over the waves
// in fixed order when the batch's slots are written. One partial per wave keeps the LDS
// footprint small enough for several waves per SIMD.
// Every slot flag is zero when the kernel starts (the key emission clears them, or a
// fill before a second backward of the same frame). A slot whose nine values are all
// zero (a record culled from the tile, behind every pixel's last contributor, or valid
// in no pixel) is not touched: its floats stay as they are, and the slot sums skip it.
// Skipping a zero term changes no sum: the sums start at +0, and adding +0 or -0 to +0
// or to a nonzero value returns it unchanged.
constexpr int NQ = 9;  // sum vs, vs dx, vs dy, vs dx^2, vs dx dy, vs dy^2, fac vr, fac vg, fac vb

// With CHUNK the workgroup runs chunk ch.k of the tile's list (a depth-chunked forward
// made the frame): each pixel starts from the chunk's end transmittance and from S over
// the final transmittance's background plus the later chunks' colours.
// With AUX (rasterize_backward with f.aux, cpu:963-1025): each pixel also carries the
// prior losses' gradients v_d = dL/dD and v_a = dL/dA, which add
// (z T - Sd / (1 - alpha)) v_d + (T - Sa / (1 - alpha)) v_a to v_alpha, Sd and Sa being
// the sums of z fac and fac over the records behind. These terms have the colour term's
// form T c - S / (1 - alpha) with c = z v_d + v_a, so one running sum S of fac (c_rgb +
// z v_d + v_a) carries all three; the per-pixel background is bg_img, and each record's
// sum of fac v_d goes to its depth slot.
template <int PPT, int BATCH, bool CHUNK = false, bool AUX = false>
__device__ inline void backward_tile(const TileView &v, const TileData &d, int gt, int lt,
                                     const ChunkView *ch = nullptr) {
#pragma clang fp contract(fast)
    constexpr int NT = 256 / PPT, NW = NT / 64;
    __shared__ float4 la[BATCH], lb[BATCH];
    __shared__ float4 lconic[BATCH];  // the record's conic as projected (la and lb hold it scaled)
    __shared__ float lc[BATCH];
    __shared__ float lz[AUX ? BATCH : 1];             // the record's view depth
    __shared__ float partd[AUX ? NW : 1][AUX ? BATCH : 1];  // per-wave sums of fac v_d
    __shared__ uint32_t lslot[BATCH];
    __shared__ int lk[BATCH];       // the staged record's position in the tile's list
    __shared__ uint8_t lmask[BATCH];  // bit c: raw colour c + 0.5 >= 0 (its gradient passes)
    __shared__ uint8_t ltouch[BATCH];  // some wave had a valid pair with the record
    __shared__ float part[NW][BATCH][NQ];
    __shared__ int smax[NW];
    __shared__ int wcount[2][NW];
    const int tx = lt % v.tiles_x, ty = lt / v.tiles_x;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int s = d.tile_start[gt], e = d.tile_start[gt + 1];
    if (CHUNK) chunk_range(s, e, ch->k, ch->K, s, e);
    if (s == e) return;
    const SumLanes sl(lane);
    const int px = tx * TILE + (lane & 15);
    const float x0 = (float)(tx * TILE), y0 = (float)(ty * TILE);
    float fy[PPT], T[PPT], S[PPT], vr[PPT], vg[PPT], vb[PPT];
    float vd[AUX ? PPT : 1], va[AUX ? PPT : 1];
    int last[PPT];
    int m = s - 1;
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const int py = ty * TILE + 4 * PPT * w + (lane >> 4) + 4 * k;
        fy[k] = (float)py;
        T[k] = 0.f; S[k] = 0.f; vr[k] = 0.f; vg[k] = 0.f; vb[k] = 0.f;
        if (AUX) { vd[k] = 0.f; va[k] = 0.f; }
        last[k] = s - 1;
        if (px < (int)v.width && py < (int)v.height) {
            const size_t pid = (size_t)py * v.width + px;
            T[k] = v.final_T[pid];
            last[k] = min(v.final_idx[pid], e - 1);
            vr[k] = v.v_image[3 * pid];
            vg[k] = v.v_image[3 * pid + 1];
            vb[k] = v.v_image[3 * pid + 2];
            if (AUX) {
                vd[k] = v.v_aux[2 * pid];
                va[k] = v.v_aux[2 * pid + 1];
                S[k] = T[k] * (v.bg_img[3 * pid] * vr[k] + v.bg_img[3 * pid + 1] * vg[k] + v.bg_img[3 * pid + 2] * vb[k]);
            } else {
                S[k] = T[k] * (v.bg[0] * vr[k] + v.bg[1] * vg[k] + v.bg[2] * vb[k]);
            }
            if (CHUNK) {  // the chunk's end transmittance and S over everything behind it
                const size_t o = (size_t)ch->k * ch->npix + pid;
                T[k] = fabsf(ch->state[o].w);
                S[k] = ch->prod[o];
            }
        }
        m = max(m, last[k]);
    }
    // the block's furthest contributing intersection: nothing behind it is visited
    for (int dd = 32; dd >= 1; dd >>= 1) m = max(m, __shfl_xor(m, dd));
    if (lane == 0) smax[w] = m;
    __syncthreads();
    int kmax = smax[0];
#pragma unroll
    for (int i = 1; i < NW; i++) kmax = max(kmax, smax[i]);
    const float fx = (float)px;

    uint32_t nwritten = 0;  // slots this thread writes (counted when slots_written is given)
    for (int end = kmax + 1; end > s; end -= BATCH) {
        const int start = max(s, end - BATCH);
        const int n = end - start;
        __syncthreads();
        // stage the batch's records that can reach the tile, in list order
        int kept = 0;
        for (int c = 0; c < n; c += NT) {
            const int i = c + threadIdx.x;
            bool keep = false;
            float4 A0, B0;
            float C0 = 0.f;
            uint32_t p = 0;
            float Z0 = 0.f;
            if (i < n) {
                const int g = (int)((uint32_t)d.gid[start + i] & d.gid_mask);
                A0 = d.splat_a[g];
                B0 = d.splat_b[g];
                C0 = d.splat_blue[g];
                if (AUX) Z0 = d.depths[g];
                keep = !(d.cull && splat_misses_tile(A0, B0, x0, y0));
                if (keep) {
                    const uint2 r = d.rect[g];
                    const int rx0 = r.x & 0xFFFF, rx1 = r.x >> 16, ry0 = r.y & 0xFFFF;
                    p = d.offsets[g] + (uint32_t)((ty - ry0) * (rx1 - rx0) + (tx - rx0));
                }
            }
            int total;
            const int pos = kept + kept_position<NW>(keep, wcount[(c / NT) & 1], total);
            if (keep) {  // colours stored clamped, max(raw + 0.5, 0), as composited
                la[pos] = staged_a(A0);
                lb[pos] = make_float4(B0.x * POW_HALF, B0.y, fmaxf(B0.z + 0.5f, 0.f), fmaxf(B0.w + 0.5f, 0.f));
                lconic[pos] = make_float4(A0.z, A0.w, B0.x, 0.f);
                lc[pos] = fmaxf(C0 + 0.5f, 0.f);
                lmask[pos] = (uint8_t)((B0.z + 0.5f >= 0.f ? 1 : 0) | (B0.w + 0.5f >= 0.f ? 2 : 0) |
                                       (C0 + 0.5f >= 0.f ? 4 : 0));
                lslot[pos] = p;
                lk[pos] = start + i;
                ltouch[pos] = 0;
                if (AUX) lz[pos] = Z0;
            }
            kept += total;
        }
        for (int i = threadIdx.x; i < NW * NQ * BATCH; i += NT) (&part[0][0][0])[i] = 0.f;
        if (AUX)
            for (int i = threadIdx.x; i < NW * BATCH; i += NT) (&partd[0][0])[i] = 0.f;
        __syncthreads();
        for (int j = kept - 1; j >= 0; j--) {
            const int kk = lk[j];
            const float4 A = la[j], B = lb[j];
            const float cr = B.z, cg = B.w, cbl = lc[j];
            // the lane's pixels share a column: dx and its products are per gaussian, and
            // the dx-weighted sums follow from q0 and q2 after the rows
            const float dx = A.x - fx;
            const float adx2 = (A.z * dx) * dx, bdx = A.w * dx;
            float q0 = 0.f, q2 = 0.f, q5 = 0.f, q6 = 0.f, q7 = 0.f, q8 = 0.f, q9 = 0.f;
            const float z = AUX ? lz[j] : 0.f;
            bool contrib = false;
#pragma unroll
            for (int k = 0; k < PPT; k++) {
                const float dy = A.y - fy[k];
                const float pw = splat_power(adx2, bdx, B.x, dy);
                const float alpha = splat_alpha(B.y, pw);
                const bool valid = kk <= last[k] && splat_valid(pw, alpha);
                contrib = contrib || valid;
                // predicated: an invalid pair leaves T and S unchanged and adds zeros
                const float ra = __builtin_amdgcn_rcpf(1.f - alpha);
                const float Tn = T[k] * ra;
                const float fac = valid ? alpha * Tn : 0.f;
                float cv = cr * vr[k] + cg * vg[k] + cbl * vb[k];
                if (AUX) {
                    cv += z * vd[k] + va[k];
                    q9 += fac * vd[k];
                }
                const float v_alpha = Tn * cv - S[k] * ra;
                T[k] = valid ? Tn : T[k];
                S[k] += fac * cv;
                q6 += fac * vr[k];
                q7 += fac * vg[k];
                q8 += fac * vb[k];
                const float vs = (valid && alpha < ALPHA_MAX) ? -alpha * v_alpha : 0.f;
                const float vsy = vs * dy;
                q0 += vs;
                q2 += vsy;
                q5 += vsy * dy;
            }
            const float q1 = q0 * dx, q3 = q1 * dx, q4 = q2 * dx;
            if (__ballot(contrib) != 0) {
                const float t = AUX ? wave_sums10(sl, q0, q1, q2, q3, q4, q5, q6, q7, q8, q9)
                                    : wave_sums9(sl, q0, q1, q2, q3, q4, q5, q6, q7, q8);
                if (lane < NQ) part[w][j][sl.q] = t;
                if (lane == 0) ltouch[j] = 1;
                if (AUX && lane == 9) partd[w][j] = t;  // lane 9: the total of q9
            }
        }
        __syncthreads();
        for (int j = threadIdx.x; j < kept; j += NT) {
            const uint32_t p = lslot[j];
            if (!ltouch[j]) continue;  // every wave's partials are zero
            nwritten++;
            float q[NQ];
#pragma unroll
            for (int i = 0; i < NQ; i++) {
                float acc = part[0][j][i];
#pragma unroll
                for (int r = 1; r < NW; r++) acc += part[r][j][i];
                q[i] = acc;
            }
            const float4 Q = lconic[j];  // the unscaled conic (a, b, c)
            const float ox = Q.x * q[1] + Q.y * q[2], oy = Q.y * q[1] + Q.z * q[2];
            const uint32_t cm = lmask[j];
            const bool mr = cm & 1u, mg = cm & 2u, mb = cm & 4u;
            d.slot_a[p] = make_float4(ox, oy, 0.5f * q[3], 0.5f * q[4]);
            d.slot_b[p] = make_float4(0.5f * q[5], mr ? q[6] : 0.f, mg ? q[7] : 0.f, mb ? q[8] : 0.f);
            d.slot_c[p] = -(1.f - lb[j].y) * q[0];
            d.slot_f[p] = 1;
            if (AUX) {
                float acc = partd[0][j];
#pragma unroll
                for (int r = 1; r < NW; r++) acc += partd[r][j];
                d.slot_d[p] = acc;
            }
        }
    }
    if (d.slots_written) {  // one atomic per wave
        for (int dd = 32; dd >= 1; dd >>= 1) nwritten += __shfl_xor(nwritten, dd);
        if (lane == 0 && nwritten) atomicAdd(d.slots_written, (unsigned long long)nwritten);
    }
}

}  // namespace msplat
