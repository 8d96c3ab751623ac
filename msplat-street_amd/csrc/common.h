// Shared device helpers for the gfx950 kernels: constants of the splatting math,
// wave64 reductions over DPP, and the tile geometry.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace msplat {

constexpr int TILE = 16;           // binning tile edge (pixels)
constexpr int TILE_PIX = TILE * TILE;
constexpr int WAVE = 64;

// Real SH basis constants (values of the real spherical harmonics, as fp32)
__device__ constexpr float SH_C0 = 0.28209479177387814f;
__device__ constexpr float SH_C1 = 0.4886025119029199f;
__device__ constexpr float SH_C2_0 = 1.0925484305920792f, SH_C2_1 = -1.0925484305920792f,
                           SH_C2_2 = 0.31539156525252005f, SH_C2_3 = -1.0925484305920792f,
                           SH_C2_4 = 0.5462742152960396f;
__device__ constexpr float SH_C3_0 = -0.5900435899266435f, SH_C3_1 = 2.890611442640554f,
                           SH_C3_2 = -0.4570457994644658f, SH_C3_3 = 0.3731763325901154f,
                           SH_C3_4 = -0.4570457994644658f, SH_C3_5 = 1.445305721320277f,
                           SH_C3_6 = -0.5900435899266435f;
__device__ constexpr float SH_C4_0 = 2.5033429417967046f, SH_C4_1 = -1.7701307697799304f,
                           SH_C4_2 = 0.9461746957575601f, SH_C4_3 = -0.6690465435572892f,
                           SH_C4_4 = 0.10578554691520431f, SH_C4_5 = -0.6690465435572892f,
                           SH_C4_6 = 0.47308734787878004f, SH_C4_7 = -1.7701307697799304f,
                           SH_C4_8 = 0.6258357354491761f;

// Compositing thresholds
__device__ constexpr float SIGMA_MAX = 5.55f;
__device__ constexpr float ALPHA_MAX = 0.999f;
__device__ constexpr float ALPHA_MIN = 1.f / 255.f;
__device__ constexpr float T_MIN = 1e-4f;

__host__ __device__ inline uint32_t sh_bases(uint32_t degree) {
    return degree == 0 ? 1 : degree == 1 ? 4 : degree == 2 ? 9 : degree == 3 ? 16 : 25;
}

// exp(x) rounded once from double: agrees with a correctly rounded expf except in
// double-rounding corner cases, so thresholds computed from it (culling, keys) match
// a host libm.
__device__ inline float exp_cr(float x) { return (float)exp((double)x); }
__device__ inline float log_cr(float x) { return (float)log((double)x); }
__device__ inline float sin_cr(float x) { return (float)sin((double)x); }
__device__ inline float cos_cr(float x) { return (float)cos((double)x); }

// Tile rectangle [x0, x1) x [y0, y1) of a splat with pixel center (cx, cy) and
// half-extents (rx, ry), clamped to the grid.
__device__ inline void tile_rect(float cx, float cy, float rx, float ry, int tx, int ty, int &x0, int &y0, int &x1,
                                 int &y1) {
    float ccx = cx / (float)TILE, ccy = cy / (float)TILE, rrx = rx / (float)TILE, rry = ry / (float)TILE;
    x0 = min(max(0, (int)(ccx - rrx)), tx);
    x1 = min(max(0, (int)(ccx + rrx + 1.f)), tx);
    y0 = min(max(0, (int)(ccy - rry)), ty);
    y1 = min(max(0, (int)(ccy + rry + 1.f)), ty);
}

// ── wave64 sum over DPP ────────────────────────────────────────────────────
// Lanes of a 16-lane row are combined by quad permutes and the row (half-)mirrors,
// the four rows by row broadcasts; the total ends in lane 63 and is returned as a
// wave-uniform value.
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}

__device__ inline float wave_sum(float v) {
    v += dpp_f<0xb1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f<0x4e>(v);   // quad_perm [2,3,0,1]
    v += dpp_f<0x141>(v);  // row_half_mirror
    v += dpp_f<0x140>(v);  // row_mirror
    v += dpp_f<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Sum over the 16 lanes of each DPP row; every lane of the row ends with the row total.
__device__ inline float row_sum16(float v) {
    v += dpp_f<0xb1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f<0x4e>(v);   // quad_perm [2,3,0,1]
    v += dpp_f<0x141>(v);  // row_half_mirror
    v += dpp_f<0x140>(v);  // row_mirror
    return v;
}

// Completes a row_sum16: lane 63 ends with the total of the four rows' sums.
__device__ inline float rows_to_lane63(float v) {
    v += dpp_f<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// Lane roles of wave_sums9 within a row of 16 lanes (r = lane & 15): at butterfly step
// s (partners r^1, r^2, then the half-row and row mirrors r^7, r^15) a lane keeps the
// first value of each pair when its bit is clear and the second when set. The bits are
// b0^b2, b1^b2, b2^b3 of r: partners differ in the bit of their own step and agree in
// the bits of the steps before and after it, so both hold the same value when they add.
struct SumLanes {
    bool s1, s2, s3;
    int q;  // value whose wave total this lane holds at the end: 4 s3 + 2 s2 + s1 (r < 8), 8 (r = 8)
    __device__ explicit SumLanes(int lane) {
        const int r = lane & 15;
        s1 = ((r ^ (r >> 2)) & 1) != 0;
        s2 = (((r >> 1) ^ (r >> 2)) & 1) != 0;
        s3 = (((r >> 2) ^ (r >> 3)) & 1) != 0;
        q = r < 8 ? 4 * (int)s3 + 2 * (int)s2 + (int)s1 : 8;
    }
};

// keep + partner's copy of the kept value (the partner sends the value it does not keep)
template <int CTRL>
__device__ inline float pair_step(float a, float b, bool second) {
    const float keep = second ? b : a, send = second ? a : b;
    return keep + dpp_f<CTRL>(send);
}

// x(row) + x(row ^ 1), then + the same over rows ^ 2: every lane ends with the sum over
// the four rows of its position's value
__device__ inline float sum_rows(float x) {
    const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    const float y = __uint_as_float(a[0]) + __uint_as_float(a[1]);
    const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(y), __float_as_uint(y), false, false);
    return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}

// Wave totals of nine per-lane values: lane l < 8 returns the total of value
// SumLanes(l).q, lane 8 the total of q8. Each value is added in the order of a plain
// per-value butterfly (partners r^1, r^2, r^7, r^15, then rows r^1, r^2), with a third
// of its DPP operations.
__device__ inline float wave_sums9(const SumLanes &L, float q0, float q1, float q2, float q3, float q4, float q5,
                                   float q6, float q7, float q8) {
    const float p0 = pair_step<0xb1>(q0, q1, L.s1), p1 = pair_step<0xb1>(q2, q3, L.s1);
    const float p2 = pair_step<0xb1>(q4, q5, L.s1), p3 = pair_step<0xb1>(q6, q7, L.s1);
    float e = q8 + dpp_f<0xb1>(q8);                              // quad_perm [1,0,3,2]
    const float r0 = pair_step<0x4e>(p0, p1, L.s2), r1 = pair_step<0x4e>(p2, p3, L.s2);
    e += dpp_f<0x4e>(e);                                         // quad_perm [2,3,0,1]
    float u = pair_step<0x141>(r0, r1, L.s3);
    e += dpp_f<0x141>(e);                                        // row_half_mirror
    u += dpp_f<0x140>(u);                                        // row_mirror
    e += dpp_f<0x140>(e);
    return sum_rows(L.q == 8 ? e : u);  // lane position r holds the same value in every row
}

// wave_sums9 with a tenth value paired with q8: lanes r >= 8 return the total of q8 where
// s1 is clear (r = 8 among them) and of q9 where it is set (r = 9 among them). q0..q8
// are added exactly as in wave_sums9.
__device__ inline float wave_sums10(const SumLanes &L, float q0, float q1, float q2, float q3, float q4, float q5,
                                    float q6, float q7, float q8, float q9) {
    const float p0 = pair_step<0xb1>(q0, q1, L.s1), p1 = pair_step<0xb1>(q2, q3, L.s1);
    const float p2 = pair_step<0xb1>(q4, q5, L.s1), p3 = pair_step<0xb1>(q6, q7, L.s1);
    float e = pair_step<0xb1>(q8, q9, L.s1);  // the partners r^2, r^7 and r^15 share s1
    const float r0 = pair_step<0x4e>(p0, p1, L.s2), r1 = pair_step<0x4e>(p2, p3, L.s2);
    e += dpp_f<0x4e>(e);
    float u = pair_step<0x141>(r0, r1, L.s3);
    e += dpp_f<0x141>(e);
    u += dpp_f<0x140>(u);
    e += dpp_f<0x140>(e);
    return sum_rows(L.q == 8 ? e : u);
}

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

// Orders LDS accesses of the lanes of one wave (no instruction is emitted).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The SH coefficients (3 DC floats, then the first u - 3 rest floats) of the lanes of a
// wave with `need`, loaded together: consecutive lanes read consecutive floats of one
// gaussian's coefficients, so a load instruction touches a few cache lines instead of one
// per lane. Lane l's coefficients land at wl[l * us, l * us + u); us is odd, so each lane
// reading its own coefficients hits a different LDS bank. gw is the wave's first
// gaussian, stride the rest floats per gaussian, widx 64 bytes of LDS. Every lane of
// the wave calls it.
__device__ inline void stage_sh(bool need, int gw, const float *dc, const float *rest, size_t stride, int u, int us,
                                float *wl, uint8_t *widx) {
    const int lane = threadIdx.x & 63;
    const uint64_t m = __ballot(need);
    if (need) widx[__popcll(m & ((1ull << lane) - 1ull))] = (uint8_t)lane;
    wave_sync();
    const int total = __popcll(m) * u;
    constexpr int B = 16;  // loads in flight per lane before their LDS writes
    for (int t0 = 0; t0 < total; t0 += 64 * B) {
        float v[B];
        int at[B];
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int t = t0 + 64 * k + lane;
            at[k] = -1;
            v[k] = 0.f;
            if (t < total) {
                const int j = t / u, c = t - j * u;
                const int l = widx[j];
                const size_t gg = (size_t)(gw + l);
                v[k] = c < 3 ? dc[3 * gg + c] : rest[stride * gg + (size_t)(c - 3)];
                at[k] = l * us + c;
            }
        }
#pragma unroll
        for (int k = 0; k < B; k++)
            if (at[k] >= 0) wl[at[k]] = v[k];
    }
    wave_sync();
}

// ── counter-based RNG: PCG hash (Jarzynski & Olano 2020), 24-bit uniforms in (0, 1),
//    Box-Muller normals (the reference's pcg_hash / rand_uniform / rand_normal3) ──
__host__ __device__ inline uint32_t pcg_hash(uint32_t v) {
    const uint32_t st = v * 747796405u + 2891336453u;
    const uint32_t w = ((st >> ((st >> 28u) + 4u)) ^ st) * 277803737u;
    return (w >> 22u) ^ w;
}

__device__ inline float rand_uniform(uint32_t seed, uint32_t ctr) {
    const uint32_t h = pcg_hash(seed ^ pcg_hash(ctr));
    return ((float)(h >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

__device__ inline void rand_normal3(uint32_t seed, uint32_t stream, float *out) {
    const uint32_t c = stream * 4u;
    const float u0 = rand_uniform(seed, c), u1 = rand_uniform(seed, c + 1u), u2 = rand_uniform(seed, c + 2u),
                u3 = rand_uniform(seed, c + 3u);
    const float r0 = sqrtf(-2.0f * log_cr(u0)), r1 = sqrtf(-2.0f * log_cr(u2));
    const float tau = 6.28318530717958647692f;
    out[0] = r0 * cos_cr(tau * u1);
    out[1] = r0 * sin_cr(tau * u1);
    out[2] = r1 * cos_cr(tau * u3);
}

__global__ inline void fill_u32(uint32_t *p, uint32_t v) { *p = v; }

// p[0, n) = 0, grid-stride
__global__ inline void zero_u32(uint32_t *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0u;
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on XCD b % 8,
// each XCD with its own L2). This maps workgroup b to item t so that every XCD walks its
// own contiguous chunk of [0, n) in order: neighbouring tiles (which share gaussians and
// write neighbouring gradient slots) meet in the same L2. A bijection on [0, n).
constexpr int NUM_XCD = 8;
__host__ __device__ inline int xcd_item(int b, int n) {
    const int x = b % NUM_XCD, l = b / NUM_XCD;
    const int q = n / NUM_XCD, r = n % NUM_XCD;
    return x * q + (x < r ? x : r) + l;
}

}  // namespace msplat
