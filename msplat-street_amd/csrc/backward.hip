// Per-gaussian backward: reduction of the intersection slots, projection/EWA/SH VJPs,
// and either
//   BWD_ADAM  — Adam on all six groups fused (SH groups for visible gaussians only, as
//               project_and_sh_backward_adam, the others dense over every gaussian as
//               fused_adam; core/cpu/msplat_cpu.cpp:1068-1147, 1273-1292) plus the
//               densification statistics (cpu:1311-1324), in one pass over the
//               parameters; or
//   BWD_GRADS — the parameter gradients themselves (autograd backend).
// The slot sum runs over the gaussian's intersections in tile order, as the reference's
// reduce_isect_grads (cpu:1034-1064). Built with -ffp-contract=off.
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

template <int MODE>
__global__ __launch_bounds__(256) void backward_kernel(BackwardArgs a) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= a.n) return;
    const bool vis = a.radii[g] > 0;
    float vxy[2] = {0.f, 0.f}, vcn[3] = {0.f, 0.f, 0.f}, vrgb[3] = {0.f, 0.f, 0.f}, vop = 0.f;
    if (vis) {
        const uint32_t o = a.offsets[g], c = a.count[g];
        for (uint32_t p = o; p < o + c; p++) {
            const float4 A = a.slot_a[p], B = a.slot_b[p];
            vxy[0] += A.x; vxy[1] += A.y; vcn[0] += A.z; vcn[1] += A.w;
            vcn[2] += B.x; vrgb[0] += B.y; vrgb[1] += B.z; vrgb[2] += B.w;
            vop += a.slot_c[p];
        }
    }
    if (a.v_xy) {
        a.v_xy[2 * g] = vxy[0]; a.v_xy[2 * g + 1] = vxy[1];
        for (int i = 0; i < 3; i++) { a.v_conic[3 * g + i] = vcn[i]; a.v_rgb[3 * g + i] = vrgb[i]; }
        a.v_opac[g] = vop;
    }
    const msplat_hip_camera &cam = a.cam;
    const uint32_t nb = sh_bases(cam.degree);
    const size_t rest_base = (size_t)(nb - 1) * 3 * g;
    float vm[3] = {0.f, 0.f, 0.f}, vs[3] = {0.f, 0.f, 0.f}, vq[4] = {0.f, 0.f, 0.f, 0.f};
    float dirv[3] = {0.f, 0.f, 0.f};
    if (vis) {
        const float p[3] = {a.means[3 * g], a.means[3 * g + 1], a.means[3 * g + 2]};
        pixel_vjp(cam.projmat, p, cam.width, cam.height, vxy[0], vxy[1], vm);
        // + viewmat row 2 * v_depth, with v_depth = 0 outside the prior-guided path
        vm[0] = vm[0] + cam.viewmat[8] * 0.f;
        vm[1] = vm[1] + cam.viewmat[9] * 0.f;
        vm[2] = vm[2] + cam.viewmat[10] * 0.f;
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
    float basis[15];
    const int nbasis = vis ? sh_basis(min(cam.degrees_to_use, 3u), dirv, basis) : 0;

    if (MODE == BWD_GRADS) {
        for (int i = 0; i < 3; i++) {
            a.g_means[3 * g + i] = vm[i];
            a.g_scales[3 * g + i] = vs[i];
            a.g_dc[3 * g + i] = vis ? SH_C0 * vrgb[i] : 0.f;
        }
        for (int i = 0; i < 4; i++) a.g_quats[4 * g + i] = vq[i];
        for (uint32_t b = 0; b + 1 < nb; b++)
            for (int c = 0; c < 3; c++)
                a.g_rest[rest_base + 3 * b + c] = (int)b < nbasis ? basis[b] * vrgb[c] : 0.f;
        a.g_opac[g] = vop;
        return;
    }

    const float b1 = a.beta1, b2 = a.beta2, eps = a.eps;
    if (vis) {  // SH groups: visible gaussians only
        for (int c = 0; c < 3; c++) {
            const size_t i = 3 * (size_t)g + c;
            adam_step(a.dc[i], a.m1[3][i], a.m2[3][i], SH_C0 * vrgb[c], a.step[3], b1, b2, a.bc2[3], eps);
        }
        for (int b = 0; b < nbasis; b++)
            for (int c = 0; c < 3; c++) {
                const size_t i = rest_base + 3 * b + c;
                adam_step(a.rest[i], a.m1[4][i], a.m2[4][i], basis[b] * vrgb[c], a.step[4], b1, b2, a.bc2[4], eps);
            }
    }
    // dense groups: means, scales, quats, opacity
    for (int i = 0; i < 3; i++) {
        const size_t k = 3 * (size_t)g + i;
        adam_step(a.means_rw[k], a.m1[0][k], a.m2[0][k], vm[i], a.step[0], b1, b2, a.bc2[0], eps);
    }
    for (int i = 0; i < 3; i++) {
        const size_t k = 3 * (size_t)g + i;
        adam_step(a.scales[k], a.m1[1][k], a.m2[1][k], vs[i], a.step[1], b1, b2, a.bc2[1], eps);
    }
    for (int i = 0; i < 4; i++) {
        const size_t k = 4 * (size_t)g + i;
        adam_step(a.quats[k], a.m1[2][k], a.m2[2][k], vq[i], a.step[2], b1, b2, a.bc2[2], eps);
    }
    adam_step(a.opac[g], a.m1[5][g], a.m2[5][g], vop, a.step[5], b1, b2, a.bc2[5], eps);
    if (vis) {
        a.vis[g] += 1.0f;
        a.gn[g] += sqrtf(vxy[0] * vxy[0] + vxy[1] * vxy[1]);
        a.ms[g] = fmaxf(a.ms[g], (float)a.radii[g] * a.inv_max_dim);
    }
}

void launch_backward(const BackwardArgs &a, BackwardMode mode, hipStream_t s) {
    const int nb = (a.n + 255) / 256;
    if (nb == 0) return;
    if (mode == BWD_ADAM)
        backward_kernel<BWD_ADAM><<<nb, 256, 0, s>>>(a);
    else
        backward_kernel<BWD_GRADS><<<nb, 256, 0, s>>>(a);
}

}  // namespace msplat
