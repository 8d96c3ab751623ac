// Per-gaussian backward: reduction of the intersection slots, projection/EWA/SH VJPs,
// and either
//   BWD_ADAM  — Adam on all six groups fused (SH groups for visible gaussians only, as
//               project_and_sh_backward_adam, the others dense over every gaussian as
//               fused_adam; core/cpu/msplat_cpu.cpp:1068-1147, 1273-1292) plus the
//               densification statistics (cpu:1311-1324), in one pass over the
//               parameters; or
//   BWD_GRADS — the parameter gradients themselves (autograd backend).
// The slot sum runs over the gaussian's intersections in tile order, as the reference's
// reduce_isect_grads (cpu:1034-1064); slots flagged zero by the rasterizer are not read
// (they add exact zeros). Built with -ffp-contract=off.
#include <cstring>

#include "common.h"
#include "kernels.h"
#include "proj_math.h"

namespace msplat {

__device__ inline void adam_step(float &p, float &m1, float &m2, float g, float step, float b1, float b2, float bc2,
                                 float eps) {
    const float m = b1 * m1 + (1.0f - b1) * g;
    const float v = b2 * m2 + (1.0f - b2) * g * g;
    p -= step * m / (sqrtf(v) / bc2 + eps);
    m1 = m;
    m2 = v;
}

constexpr int BIG_REDUCE_GRID = 2048;  // 8 waves per SIMD

// The zero flags of the c <= BIG_SLOTS slots from o as a mask (bit k: slot o + k was
// written), from the at most nine aligned words that hold them. The words also hold
// bytes outside the range (other records' flags, or never-written bytes past the
// frame's last slot), so each byte is tested for nonzero on its own (bit 7 of
// ((x & 0x7F) + 0x7F) | x, no carry between bytes) before the four bits are gathered.
// The flag buffers are allocated in whole words.
__device__ inline uint64_t slot_mask(const uint8_t *sf, uint32_t o, uint32_t c) {
    if (c == 0) return 0;  // no slots (and o may be stale)
    const uint32_t *w = reinterpret_cast<const uint32_t *>(sf) + (o >> 2);
    const uint32_t nw = ((o & 3u) + c + 3u) >> 2;
    uint64_t m = 0;
#pragma unroll
    for (uint32_t j = 0; j < (BIG_SLOTS + 3 + 3) / 4; j++) {
        if (j < nw) {
            const uint32_t v = w[j];
            const uint32_t nz = ((((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v) & 0x80808080u) >> 7;
            m |= (uint64_t)((nz | (nz >> 7) | (nz >> 14) | (nz >> 21)) & 0xFu) << (4 * j);
        }
    }
    return (m >> (o & 3u)) & ((1ull << c) - 1ull);
}

// Sum of the written slots among o + (set bits of m), in increasing order, four slots'
// loads in flight at a time (padding adds zeros). The sum equals the plain in-order sum
// over every slot of the range: the others are zero.
__device__ inline void sum_slots(const float4 *sa, const float4 *sb, const float *sc, uint32_t o, uint64_t m,
                                 float v[9]) {
    while (m) {
        uint32_t k[4];
        bool h[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            h[u] = m != 0;
            k[u] = h[u] ? (uint32_t)__builtin_ctzll(m) : 0u;
            m &= m - 1ull;
        }
        float4 A[4], B[4];
        float Cs[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (h[u]) {
                A[u] = sa[o + k[u]];
                B[u] = sb[o + k[u]];
                Cs[u] = sc[o + k[u]];
            } else {
                A[u] = make_float4(0.f, 0.f, 0.f, 0.f);
                B[u] = A[u];
                Cs[u] = 0.f;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            v[0] += A[u].x; v[1] += A[u].y; v[2] += A[u].z; v[3] += A[u].w;
            v[4] += B[u].x; v[5] += B[u].y; v[6] += B[u].z; v[7] += B[u].w;
            v[8] += Cs[u];
        }
    }
}

// Sum of the written depth slots among o + (set bits of m), in increasing order
__device__ inline float sum_slot_d(const float *sd, uint32_t o, uint64_t m) {
    float v = 0.f;
    while (m) {
        v += sd[o + (uint32_t)__builtin_ctzll(m)];
        m &= m - 1ull;
    }
    return v;
}

// SP: the camera and Adam step sizes come from the device StepParams block (a captured
// step) rather than the kernel arguments; a compile-time choice, so that the kernel
// arguments stay scalar loads in the launch-by-launch step
template <int MODE, bool SP>
__global__ __launch_bounds__(256) void backward_kernel(BackwardArgs a) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= a.n) return;
    if (MODE == BWD_ADAM && a.overflow && *a.overflow) return;  // over capacity: nothing changes
    // the dense Adam groups' parameters and moments, loaded before the slot sums so that
    // their loads are in flight during the visible path's dependent loads
    float dp[11], d1[11], d2[11];
    if (MODE == BWD_ADAM) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            dp[i] = a.means_rw[3 * (size_t)g + i];
            d1[i] = a.m1[0][3 * (size_t)g + i];
            d2[i] = a.m2[0][3 * (size_t)g + i];
            dp[3 + i] = a.scales[3 * (size_t)g + i];
            d1[3 + i] = a.m1[1][3 * (size_t)g + i];
            d2[3 + i] = a.m2[1][3 * (size_t)g + i];
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            dp[6 + i] = a.quats[4 * (size_t)g + i];
            d1[6 + i] = a.m1[2][4 * (size_t)g + i];
            d2[6 + i] = a.m2[2][4 * (size_t)g + i];
        }
        dp[10] = a.opac[g];
        d1[10] = a.m1[5][g];
        d2[10] = a.m2[5][g];
    }
    const bool vis = a.radii[g] > 0;
    float vxy[2] = {0.f, 0.f}, vcn[3] = {0.f, 0.f, 0.f}, vrgb[3] = {0.f, 0.f, 0.f}, vop = 0.f, vdep = 0.f;
    if (vis && a.count[g] > BIG_SLOTS) {  // reduced by big_reduce_kernel
        const float4 A = a.big_a[g], B = a.big_b[g];
        vxy[0] = A.x; vxy[1] = A.y; vcn[0] = A.z; vcn[1] = A.w;
        vcn[2] = B.x; vrgb[0] = B.y; vrgb[1] = B.z; vrgb[2] = B.w;
        vop = a.big_c[g];
        if (a.slot_d) vdep = a.big_d[g];
    } else if (vis) {  // the written slots, in tile order
        const uint32_t o = a.offsets[g];
        float v[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        const uint64_t m = slot_mask(a.slot_f, o, a.count[g]);
        sum_slots(a.slot_a, a.slot_b, a.slot_c, o, m, v);
        vxy[0] = v[0]; vxy[1] = v[1]; vcn[0] = v[2]; vcn[1] = v[3];
        vcn[2] = v[4]; vrgb[0] = v[5]; vrgb[1] = v[6]; vrgb[2] = v[7];
        vop = v[8];
        if (a.slot_d) vdep = sum_slot_d(a.slot_d, o, m);
    }
    if (a.v_xy) {
        a.v_xy[2 * g] = vxy[0]; a.v_xy[2 * g + 1] = vxy[1];
        for (int i = 0; i < 3; i++) { a.v_conic[3 * g + i] = vcn[i]; a.v_rgb[3 * g + i] = vrgb[i]; }
        a.v_opac[g] = vop;
    }
    const msplat_hip_camera &cam = SP ? a.sp->cam : a.cam;
    float vm[3] = {0.f, 0.f, 0.f}, vs[3] = {0.f, 0.f, 0.f}, vq[4] = {0.f, 0.f, 0.f, 0.f};
    float dirv[3] = {0.f, 0.f, 0.f};
    if (vis) {
        const float p[3] = {a.means[3 * g], a.means[3 * g + 1], a.means[3 * g + 2]};
        pixel_vjp(cam.projmat, p, cam.width, cam.height, vxy[0], vxy[1], vm);
        // + viewmat row 2 * v_depth, with v_depth = 0 outside the prior-guided path
        vm[0] = vm[0] + cam.viewmat[8] * vdep;
        vm[1] = vm[1] + cam.viewmat[9] * vdep;
        vm[2] = vm[2] + cam.viewmat[10] * vdep;
        const float4 SA = a.splat_a[g], SB = a.splat_b[g];
        const float cn[3] = {SA.z, SA.w, SB.x};
        float vc2[3];
        conic_vjp(cn, vcn, vc2);
        const float es[3] = {exp_cr(a.scales[3 * g]), exp_cr(a.scales[3 * g + 1]), exp_cr(a.scales[3 * g + 2])};
        const float q[4] = {a.quats[4 * g], a.quats[4 * g + 1], a.quats[4 * g + 2], a.quats[4 * g + 3]};
        float cov[6], pv[3], vcov[6];
        covariance3d(es, cam.glob_scale, q, cov);
        const float tfx = 0.5f * (float)cam.width / cam.fx, tfy = 0.5f * (float)cam.height / cam.fy;
        cam_space(cam.viewmat, p, pv);
        ewa_vjp(cov, cam.viewmat, cam.fx, cam.fy, tfx, tfy, vc2, pv, vm, vcov);
        cov3d_vjp(es, cam.glob_scale, q, vcov, vs, vq);
        vs[0] *= es[0]; vs[1] *= es[1]; vs[2] *= es[2];
        view_direction(p, cam.cam_pos, dirv);
    }
    // the higher SH bands are handled per coefficient by sh_rest_kernel (every gaussian,
    // gradient mode) or sh_rest_adam_kernel (the visible list only, Adam mode)
    if (MODE == BWD_GRADS || vis) {
        a.sh_dir[g] = make_float4(dirv[0], dirv[1], dirv[2], vis ? 1.f : 0.f);
        a.sh_vrgb[g] = make_float4(vrgb[0], vrgb[1], vrgb[2], 0.f);
    }

    if (MODE == BWD_GRADS) {
        for (int i = 0; i < 3; i++) {
            a.g_means[3 * g + i] = vm[i];
            a.g_scales[3 * g + i] = vs[i];
            a.g_dc[3 * g + i] = vis ? SH_C0 * vrgb[i] : 0.f;
        }
        for (int i = 0; i < 4; i++) a.g_quats[4 * g + i] = vq[i];
        a.g_opac[g] = vop;
        return;
    }

    const float b1 = a.beta1, b2 = a.beta2, eps = a.eps;
    const float *stp = SP ? a.sp->step : a.step, *bc2 = SP ? a.sp->bc2 : a.bc2;
    if (vis) {  // SH DC: visible gaussians only
        for (int c = 0; c < 3; c++) {
            const size_t i = 3 * (size_t)g + c;
            adam_step(a.dc[i], a.m1[3][i], a.m2[3][i], SH_C0 * vrgb[c], stp[3], b1, b2, bc2[3], eps);
        }
    }
    // dense groups: means, scales, quats, opacity (from the values loaded at the start)
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const size_t k = 3 * (size_t)g + i;
        adam_step(dp[i], d1[i], d2[i], vm[i], stp[0], b1, b2, bc2[0], eps);
        a.means_rw[k] = dp[i];
        a.m1[0][k] = d1[i];
        a.m2[0][k] = d2[i];
        adam_step(dp[3 + i], d1[3 + i], d2[3 + i], vs[i], stp[1], b1, b2, bc2[1], eps);
        a.scales[k] = dp[3 + i];
        a.m1[1][k] = d1[3 + i];
        a.m2[1][k] = d2[3 + i];
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const size_t k = 4 * (size_t)g + i;
        adam_step(dp[6 + i], d1[6 + i], d2[6 + i], vq[i], stp[2], b1, b2, bc2[2], eps);
        a.quats[k] = dp[6 + i];
        a.m1[2][k] = d1[6 + i];
        a.m2[2][k] = d2[6 + i];
    }
    adam_step(dp[10], d1[10], d2[10], vop, stp[5], b1, b2, bc2[5], eps);
    a.opac[g] = dp[10];
    a.m1[5][g] = d1[10];
    a.m2[5][g] = d2[10];
    if (vis) {
        a.vis[g] += 1.0f;
        a.gn[g] += sqrtf(vxy[0] * vxy[0] + vxy[1] * vxy[1]);
        a.ms[g] = fmaxf(a.ms[g], (float)a.radii[g] * a.inv_max_dim);
    }
}

// Slot sums of the gaussians with more than BIG_SLOTS intersections, one wave each
// (grid-stride over the device-side list; every wave exits when the list is done).
// Lane l adds the written slots among l, l + 64, ... in order, in rounds of four slots
// whose flag and value loads are in flight together; the nine lane sums (ten with depth
// slots) are added over the wave by wave_sums9 (wave_sums10). The next record's id and slot range are loaded while this one
// is summed.
__global__ __launch_bounds__(256) void big_reduce_kernel(BackwardArgs a) {
    const uint32_t nbig = *a.big_count;
    const int lane = threadIdx.x & 63;
    const SumLanes sl(lane);
    const uint32_t wave = (blockIdx.x * 256 + threadIdx.x) >> 6, nwaves = (gridDim.x * 256) >> 6;
    uint32_t g = wave < nbig ? a.big_list[wave] : 0u;
    uint32_t o = 0, end = 0;
    if (wave < nbig) {
        o = a.offsets[g];
        end = o + a.count[g];
    }
    for (uint32_t i = wave; i < nbig; i += nwaves) {
        // the next record's id and slot range, loaded while this one is summed
        const uint32_t gn = i + nwaves < nbig ? a.big_list[i + nwaves] : 0u;
        uint32_t on = 0, cn = 0;
        if (i + nwaves < nbig) {
            on = a.offsets[gn];
            cn = a.count[gn];
        }
        float v[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, vd = 0.f;
        for (uint32_t p0 = o + lane; p0 < end; p0 += 256) {
            bool f[4];
#pragma unroll
            for (int u = 0; u < 4; u++) f[u] = p0 + 64 * u < end && a.slot_f[p0 + 64 * u] != 0;
            if (a.slot_d)
#pragma unroll
                for (int u = 0; u < 4; u++) vd += f[u] ? a.slot_d[p0 + 64 * u] : 0.f;
            float4 A[4], B[4];
            float Cs[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (f[u]) {
                    A[u] = a.slot_a[p0 + 64 * u];
                    B[u] = a.slot_b[p0 + 64 * u];
                    Cs[u] = a.slot_c[p0 + 64 * u];
                } else {
                    A[u] = make_float4(0.f, 0.f, 0.f, 0.f);
                    B[u] = A[u];
                    Cs[u] = 0.f;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                v[0] += A[u].x; v[1] += A[u].y; v[2] += A[u].z; v[3] += A[u].w;
                v[4] += B[u].x; v[5] += B[u].y; v[6] += B[u].z; v[7] += B[u].w;
                v[8] += Cs[u];
            }
        }
        // with depth slots the tenth sum rides with the ninth: lane 9 ends with its total
        const float t = a.slot_d ? wave_sums10(sl, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], vd)
                                 : wave_sums9(sl, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8]);
        if (a.slot_d && lane == 9) a.big_d[g] = t;
        if (lane < 9) {
            float *dst = sl.q < 4 ? &a.big_a[g].x + sl.q : sl.q < 8 ? &a.big_b[g].x + (sl.q - 4) : &a.big_c[g];
            *dst = t;
        }
        g = gn;
        o = on;
        end = on + cn;
    }
}

// SH bands 1..3, one thread per stored coefficient (gaussian g, basis b, channel c), so
// that the (N, bases-1, 3) coefficients and moments are read and written coalesced.
// Adam mode updates coefficients of visible gaussians in the bands in use (the
// reference's fused SH Adam, cpu:1118-1137); gradient mode writes every coefficient's
// gradient (zero outside the bands in use).
template <int MODE>
__global__ __launch_bounds__(256) void sh_rest_kernel(BackwardArgs a) {
    const uint32_t nb = sh_bases(a.cam.degree);
    const uint32_t stride = (nb - 1) * 3;
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (stride == 0 || e >= (size_t)a.n * stride) return;
    const int g = (int)(e / stride);
    const int j = (int)(e % stride), b = j / 3, c = j % 3;
    const float4 D = a.sh_dir[g];
    const int dtu = (int)min(a.cam.degrees_to_use, 3u);
    const int nbasis = dtu >= 3 ? 15 : dtu == 2 ? 8 : dtu == 1 ? 3 : 0;
    const bool active = D.w > 0.f && b < nbasis;
    if (MODE == BWD_GRADS && !active) {
        a.g_rest[e] = 0.f;
        return;
    }
    if (!active) return;
    const float d[3] = {D.x, D.y, D.z};
    float basis[15];
    sh_basis((uint32_t)dtu, d, basis);
    const float4 V = a.sh_vrgb[g];
    const float vc = c == 0 ? V.x : c == 1 ? V.y : V.z;
    const float grad = basis[b] * vc;
    if (MODE == BWD_GRADS) {
        a.g_rest[e] = grad;
        return;
    }
    adam_step(a.rest[e], a.m1[4][e], a.m2[4][e], grad, a.step[4], a.beta1, a.beta2, a.bc2[4], a.eps);
}

// Adam on SH bands 1..3 of the visible gaussians only: one thread per coefficient of
// the frame's visible list (grid-stride over its device-side count), so the many
// gaussians outside the view, whose coefficients the fused SH Adam leaves alone, cost
// no threads. A gaussian's coefficients are contiguous, so its threads stay coalesced.
template <bool SP>
__global__ __launch_bounds__(256) void sh_rest_adam_kernel(BackwardArgs a) {
    if (a.overflow && *a.overflow) return;  // over capacity: nothing changes
    const msplat_hip_camera &cam = SP ? a.sp->cam : a.cam;
    const float step = SP ? a.sp->step[4] : a.step[4], bc2 = SP ? a.sp->bc2[4] : a.bc2[4];
    const uint32_t stride = (sh_bases(cam.degree) - 1) * 3;
    const int dtu = (int)min(cam.degrees_to_use, 3u);
    const uint32_t nbasis = dtu >= 3 ? 15 : dtu == 2 ? 8 : dtu == 1 ? 3 : 0;
    const size_t total = (size_t)*a.vis_count * stride;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const uint32_t i = (uint32_t)(e / stride), j = (uint32_t)(e % stride), b = j / 3, c = j % 3;
        if (b >= nbasis) continue;
        const int g = a.vis_list[i];
        const float4 D = a.sh_dir[g];
        const float d[3] = {D.x, D.y, D.z};
        float basis[15];
        sh_basis((uint32_t)dtu, d, basis);
        const float4 V = a.sh_vrgb[g];
        const float vc = c == 0 ? V.x : c == 1 ? V.y : V.z;
        const size_t k = (size_t)g * stride + j;
        adam_step(a.rest[k], a.m1[4][k], a.m2[4][k], basis[b] * vc, step, a.beta1, a.beta2, bc2, a.eps);
    }
}

constexpr int SH_ADAM_GRID = 4096;

// ── batched views ────────────────────────────────────────────────────────────
// Gaussian g's gradient summed over the views of a batch, one thread per gaussian. Per
// view the slot sums (or the big-list sums) give the 2D cotangents, which go through
// the view's pixel, conic and EWA VJPs; the world-covariance cotangents are summed over
// the views and taken through the covariance VJP once (it is linear in them), and the
// SH band cotangents basis(dir_v) x v_rgb_v are summed per coefficient. The
// densification statistics are updated per view as a sequence of one-view steps would.
__global__ __launch_bounds__(256) void batch_backward_kernel(BatchBackwardArgs a) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= a.n || a.sc->overflow) return;
    const float p[3] = {a.means[3 * g], a.means[3 * g + 1], a.means[3 * g + 2]};
    const float es[3] = {exp_cr(a.scales[3 * g]), exp_cr(a.scales[3 * g + 1]), exp_cr(a.scales[3 * g + 2])};
    const float q[4] = {a.quats[4 * g], a.quats[4 * g + 1], a.quats[4 * g + 2], a.quats[4 * g + 3]};
    const float glob = a.tab->views[0].cam.glob_scale;
    float cov[6];
    covariance3d(es, glob, q, cov);
    float vm[3] = {0.f, 0.f, 0.f}, vcov[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, vrgb[3] = {0.f, 0.f, 0.f}, vop = 0.f;
    float racc[45];
#pragma unroll
    for (int i = 0; i < 45; i++) racc[i] = 0.f;
    bool any = false;
    float st_vis = 0.f, st_gn = 0.f, st_ms = a.ms ? a.ms[g] : 0.f;
    for (int v = 0; v < a.num_views; v++) {
        const size_t r = (size_t)v * a.n + g;
        const int32_t rad = a.radii[r];
        if (rad <= 0) continue;
        any = true;
        float vxy[2] = {0.f, 0.f}, vcn[3] = {0.f, 0.f, 0.f}, vc[3] = {0.f, 0.f, 0.f}, vo = 0.f;
        const uint32_t c = a.count[r];
        if (c > BIG_SLOTS) {
            const float4 A = a.big_a[r], B = a.big_b[r];
            vxy[0] = A.x; vxy[1] = A.y; vcn[0] = A.z; vcn[1] = A.w;
            vcn[2] = B.x; vc[0] = B.y; vc[1] = B.z; vc[2] = B.w;
            vo = a.big_c[r];
        } else {
            const uint32_t o = a.offsets[r];
            float s9[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            sum_slots(a.slot_a, a.slot_b, a.slot_c, o, slot_mask(a.slot_f, o, c), s9);
            vxy[0] = s9[0]; vxy[1] = s9[1]; vcn[0] = s9[2]; vcn[1] = s9[3];
            vcn[2] = s9[4]; vc[0] = s9[5]; vc[1] = s9[6]; vc[2] = s9[7];
            vo = s9[8];
        }
        const BatchView &bv = a.tab->views[v];
        const msplat_hip_camera &cam = bv.cam;
        float vmv[3];
        pixel_vjp(cam.projmat, p, cam.width, cam.height, vxy[0], vxy[1], vmv);
        const float4 SA = a.splat_a[r], SB = a.splat_b[r];
        const float cn[3] = {SA.z, SA.w, SB.x};
        float vc2[3], pv[3], vcv[6];
        conic_vjp(cn, vcn, vc2);
        const float tfx = 0.5f * (float)cam.width / cam.fx, tfy = 0.5f * (float)cam.height / cam.fy;
        cam_space(cam.viewmat, p, pv);
        ewa_vjp(cov, cam.viewmat, cam.fx, cam.fy, tfx, tfy, vc2, pv, vmv, vcv);
#pragma unroll
        for (int i = 0; i < 3; i++) vm[i] += vmv[i];
#pragma unroll
        for (int i = 0; i < 6; i++) vcov[i] += vcv[i];
#pragma unroll
        for (int i = 0; i < 3; i++) vrgb[i] += vc[i];
        vop += vo;
        const int dtu = (int)min(cam.degrees_to_use, 3u);
        if (dtu > 0) {
            float d[3], basis[15];
            view_direction(p, cam.cam_pos, d);
            const int nbasis = sh_basis((uint32_t)dtu, d, basis);
#pragma unroll
            for (int b = 0; b < 15; b++)
                if (b < nbasis) {
                    racc[3 * b] += basis[b] * vc[0];
                    racc[3 * b + 1] += basis[b] * vc[1];
                    racc[3 * b + 2] += basis[b] * vc[2];
                }
        }
        st_vis += 1.0f;
        st_gn += sqrtf(vxy[0] * vxy[0] + vxy[1] * vxy[1]);
        st_ms = fmaxf(st_ms, (float)rad * bv.inv_max_dim);
    }
    float vs[3] = {0.f, 0.f, 0.f}, vq[4] = {0.f, 0.f, 0.f, 0.f};
    if (any) {
        cov3d_vjp(es, glob, q, vcov, vs, vq);
        vs[0] *= es[0]; vs[1] *= es[1]; vs[2] *= es[2];
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        a.g_means[3 * g + i] = vm[i];
        a.g_scales[3 * g + i] = vs[i];
        a.g_dc[3 * g + i] = any ? SH_C0 * vrgb[i] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) a.g_quats[4 * g + i] = vq[i];
    a.g_opac[g] = vop;
    if (a.rest_stride) {
        float *gr = a.g_rest + (size_t)a.rest_stride * g;
#pragma unroll
        for (int i = 0; i < 45; i++)
            if ((uint32_t)i < a.rest_stride) gr[i] = racc[i];
        for (uint32_t i = 45; i < a.rest_stride; i++) gr[i] = 0.f;
    }
    a.vis_any[g] = any ? 1 : 0;
    if (a.vis && any) {
        a.vis[g] += st_vis;
        a.gn[g] += st_gn;
        a.ms[g] = st_ms;
    }
}

__global__ __launch_bounds__(256) void batch_adam_kernel(float *p, float *m1, float *m2, const float *grad,
                                                        size_t count, int width, const uint8_t *vis,
                                                        const BatchTable *tab, int group, const uint32_t *overflow) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= count || *overflow) return;
    if (vis && !vis[e / (size_t)width]) return;
    if (group == 4 && (uint32_t)(e % (size_t)width) >= tab->rest_active) return;
    adam_step(p[e], m1[e], m2[e], grad[e], tab->step[group], tab->beta1, tab->beta2, tab->bc2[group], tab->eps);
}

void launch_big_reduce_records(const uint32_t *big_list, const uint32_t *big_count, const uint32_t *offsets,
                               const uint32_t *count, const float4 *slot_a, const float4 *slot_b, const float *slot_c,
                               const uint8_t *slot_f, float4 *big_a, float4 *big_b, float *big_c, hipStream_t s) {
    BackwardArgs a;
    memset(&a, 0, sizeof a);
    a.big_list = big_list;
    a.big_count = big_count;
    a.offsets = offsets;
    a.count = count;
    a.slot_a = slot_a;
    a.slot_b = slot_b;
    a.slot_c = slot_c;
    a.slot_f = slot_f;
    a.big_a = big_a;
    a.big_b = big_b;
    a.big_c = big_c;
    big_reduce_kernel<<<BIG_REDUCE_GRID, 256, 0, s>>>(a);
}

void launch_batch_backward(const BatchBackwardArgs &a, hipStream_t s) {
    const int nb = (a.n + 255) / 256;
    if (nb > 0 && a.num_views > 0) batch_backward_kernel<<<nb, 256, 0, s>>>(a);
}

void launch_batch_adam(float *p, float *m1, float *m2, const float *grad, size_t count, int width,
                       const uint8_t *vis, const BatchTable *tab, int group, const uint32_t *overflow, hipStream_t s) {
    const size_t nb = (count + 255) / 256;
    if (nb > 0) batch_adam_kernel<<<(unsigned)nb, 256, 0, s>>>(p, m1, m2, grad, count, width, vis, tab, group, overflow);
}

void launch_backward(const BackwardArgs &a, BackwardMode mode, hipStream_t s) {
    const int nb = (a.n + 255) / 256;
    if (nb == 0) return;
    const size_t ne = (size_t)a.n * (sh_bases(a.cam.degree) - 1) * 3;
    const int nbe = (int)((ne + 255) / 256);
    big_reduce_kernel<<<BIG_REDUCE_GRID, 256, 0, s>>>(a);
    if (mode == BWD_ADAM) {
        if (a.sp) {
            backward_kernel<BWD_ADAM, true><<<nb, 256, 0, s>>>(a);
            if (nbe) sh_rest_adam_kernel<true><<<std::min(nbe, SH_ADAM_GRID), 256, 0, s>>>(a);
        } else {
            backward_kernel<BWD_ADAM, false><<<nb, 256, 0, s>>>(a);
            if (nbe) sh_rest_adam_kernel<false><<<std::min(nbe, SH_ADAM_GRID), 256, 0, s>>>(a);
        }
    } else {
        backward_kernel<BWD_GRADS, false><<<nb, 256, 0, s>>>(a);
        if (nbe) sh_rest_kernel<BWD_GRADS><<<nbe, 256, 0, s>>>(a);
    }
}

}  // namespace msplat
