// Projection math shared by the forward and backward kernels: EWA splatting of 3D
// Gaussians, SH colour, and their vector-Jacobian products. Written for exact IEEE
// single precision (the library builds with -ffp-contract=off), in the operation order
// of the reference's CPU backend (core/cpu/msplat_cpu.cpp:63-291), so that projection
// outputs, radii and tile rectangles agree bit for bit with a host evaluation.
#pragma once

#include "common.h"

namespace msplat {

struct m3 {
    float a[3][3];  // a[col][row], column-major like the reference's MSL types
};

__device__ inline m3 mm(const m3 &A, const m3 &B) {
    m3 R;
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) R.a[j][i] = A.a[0][i] * B.a[j][0] + A.a[1][i] * B.a[j][1] + A.a[2][i] * B.a[j][2];
    return R;
}

__device__ inline m3 tr(const m3 &A) {
    m3 R;
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) R.a[j][i] = A.a[i][j];
    return R;
}

__device__ inline m3 cols(float a0, float a1, float a2, float b0, float b1, float b2, float c0, float c1, float c2) {
    m3 R;
    R.a[0][0] = a0; R.a[0][1] = a1; R.a[0][2] = a2;
    R.a[1][0] = b0; R.a[1][1] = b1; R.a[1][2] = b2;
    R.a[2][0] = c0; R.a[2][1] = c1; R.a[2][2] = c2;
    return R;
}

__device__ inline float dot3(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Rotation of the stored quaternion q = (w, x, y, z), normalized
__device__ inline m3 rotation(const float *q) {
    float s = 1.f / sqrtf(q[3] * q[3] + q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    float w = q[0] * s, x = q[1] * s, y = q[2] * s, z = q[3] * s;
    return cols(1.f - 2.f * (y * y + z * z), 2.f * (x * y + w * z), 2.f * (x * z - w * y),
                2.f * (x * y - w * z), 1.f - 2.f * (x * x + z * z), 2.f * (y * z + w * x),
                2.f * (x * z + w * y), 2.f * (y * z - w * x), 1.f - 2.f * (x * x + y * y));
}

__device__ inline m3 diag(const float *s, float g) { return cols(g * s[0], 0, 0, 0, g * s[1], 0, 0, 0, g * s[2]); }

// upper triangle of (R S)(R S)^T
__device__ inline void covariance3d(const float *s, float glob, const float *q, float *cov) {
    m3 M = mm(rotation(q), diag(s, glob));
    m3 C = mm(M, tr(M));
    cov[0] = C.a[0][0]; cov[1] = C.a[0][1]; cov[2] = C.a[0][2];
    cov[3] = C.a[1][1]; cov[4] = C.a[1][2]; cov[5] = C.a[2][2];
}

__device__ inline void fov_clamp(float *p, float tfx, float tfy) {
    float lx = 1.3f * tfx, ly = 1.3f * tfy;
    p[0] = p[2] * fminf(lx, fmaxf(-lx, p[0] / p[2]));
    p[1] = p[2] * fminf(ly, fmaxf(-ly, p[1] / p[2]));
}

// EWA 2D covariance (+0.3 px low-pass) of a camera-space point
__device__ inline void ewa_cov2d(const float *cov, const float *vm, float fx, float fy, float tfx, float tfy,
                                 const float *pv_in, float *c2) {
    float pv[3] = {pv_in[0], pv_in[1], pv_in[2]};
    fov_clamp(pv, tfx, tfy);
    float rz = 1.f / pv[2], rz2 = rz * rz;
    float j00 = fx * rz, j11 = fy * rz, j20 = -fx * pv[0] * rz2, j21 = -fy * pv[1] * rz2;
    float t0[3] = {vm[0] * j00 + vm[8] * j20, vm[1] * j00 + vm[9] * j20, vm[2] * j00 + vm[10] * j20};
    float t1[3] = {vm[4] * j11 + vm[8] * j21, vm[5] * j11 + vm[9] * j21, vm[6] * j11 + vm[10] * j21};
    float a = cov[0], b = cov[1], c = cov[2], d = cov[3], e = cov[4], f = cov[5];
    float u0[3] = {t0[0] * a + t0[1] * b + t0[2] * c, t0[0] * b + t0[1] * d + t0[2] * e,
                   t0[0] * c + t0[1] * e + t0[2] * f};
    float u1[3] = {t1[0] * a + t1[1] * b + t1[2] * c, t1[0] * b + t1[1] * d + t1[2] * e,
                   t1[0] * c + t1[1] * e + t1[2] * f};
    c2[0] = dot3(u0, t0) + 0.3f;
    c2[1] = dot3(u0, t1);
    c2[2] = dot3(u1, t1) + 0.3f;
}

__device__ inline bool conic_radius(const float *c2, float *conic, float &radius) {
    float det = c2[0] * c2[2] - c2[1] * c2[1];
    if (det == 0.f) return false;
    float inv = 1.f / det;
    conic[0] = c2[2] * inv;
    conic[1] = -c2[1] * inv;
    conic[2] = c2[0] * inv;
    float mid = 0.5f * (c2[0] + c2[2]);
    float disc = sqrtf(fmaxf(0.1f, mid * mid - det));
    radius = ceilf(3.f * sqrtf(mid + disc));
    return true;
}

__device__ inline void cam_space(const float *vm, const float *p, float *o) {
    o[0] = vm[0] * p[0] + vm[1] * p[1] + vm[2] * p[2] + vm[3];
    o[1] = vm[4] * p[0] + vm[5] * p[1] + vm[6] * p[2] + vm[7];
    o[2] = vm[8] * p[0] + vm[9] * p[1] + vm[10] * p[2] + vm[11];
}

__device__ inline void to_pixel(const float *pm, const float *p, uint32_t W, uint32_t H, float cx, float cy, float &ox,
                                float &oy) {
    float hx = pm[0] * p[0] + pm[1] * p[1] + pm[2] * p[2] + pm[3];
    float hy = pm[4] * p[0] + pm[5] * p[1] + pm[6] * p[2] + pm[7];
    float hw = pm[12] * p[0] + pm[13] * p[1] + pm[14] * p[2] + pm[15];
    float rw = 1.f / (hw + 1e-6f);
    ox = 0.5f * (float)(int)W * (hx * rw) + cx - 0.5f;
    oy = 0.5f * (float)(int)H * (hy * rw) + cy - 0.5f;
}

__device__ inline void view_direction(const float *p, const float *cam, float *d) {
    float v[3] = {p[0] - cam[0], p[1] - cam[1], p[2] - cam[2]};
    float s = 1.f / sqrtf(dot3(v, v));
    d[0] = v[0] * s; d[1] = v[1] * s; d[2] = v[2] * s;
}

// Raw view-dependent colour (no +0.5) from dc (3) and rest (bases x 3, channel fastest)
__device__ inline void sh_color(uint32_t deg, const float *d, const float *dc, const float *r0, float *col) {
#pragma unroll
    for (int c = 0; c < 3; c++) col[c] = SH_C0 * dc[c];
    if (deg < 1) return;
    const float x = d[0], y = d[1], z = d[2];
    const float xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float *r = r0 + c;
        col[c] += SH_C1 * (-y * r[0] + z * r[3] - x * r[6]);
        if (deg < 2) continue;
        col[c] += (SH_C2_0 * xy * r[9] + SH_C2_1 * yz * r[12] + SH_C2_2 * (2.f * zz - xx - yy) * r[15] +
                   SH_C2_3 * xz * r[18] + SH_C2_4 * (xx - yy) * r[21]);
        if (deg < 3) continue;
        col[c] += (SH_C3_0 * y * (3.f * xx - yy) * r[24] + SH_C3_1 * xy * z * r[27] +
                   SH_C3_2 * y * (4.f * zz - xx - yy) * r[30] + SH_C3_3 * z * (2.f * zz - 3.f * xx - 3.f * yy) * r[33] +
                   SH_C3_4 * x * (4.f * zz - xx - yy) * r[36] + SH_C3_5 * z * (xx - yy) * r[39] +
                   SH_C3_6 * x * (xx - 3.f * yy) * r[42]);
        if (deg < 4) continue;
        col[c] += (SH_C4_0 * xy * (xx - yy) * r[45] + SH_C4_1 * yz * (3.f * xx - yy) * r[48] +
                   SH_C4_2 * xy * (7.f * zz - 1.f) * r[51] + SH_C4_3 * yz * (7.f * zz - 3.f) * r[54] +
                   SH_C4_4 * (zz * (35.f * zz - 30.f) + 3.f) * r[57] + SH_C4_5 * xz * (7.f * zz - 3.f) * r[60] +
                   SH_C4_6 * (xx - yy) * (7.f * zz - 1.f) * r[63] + SH_C4_7 * xz * (xx - 3.f * yy) * r[66] +
                   SH_C4_8 * (xx * (xx - 3.f * yy) - yy * (3.f * xx - yy)) * r[69]);
    }
}

// Basis values of bands 1..3 in storage order (the SH backward's per-coefficient factor)
__device__ inline int sh_basis(uint32_t deg, const float *d, float *o) {
    const float x = d[0], y = d[1], z = d[2];
    const float xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    int n = 0;
    if (deg < 1) return n;
    o[n++] = -SH_C1 * y; o[n++] = SH_C1 * z; o[n++] = -SH_C1 * x;
    if (deg < 2) return n;
    o[n++] = SH_C2_0 * xy; o[n++] = SH_C2_1 * yz; o[n++] = SH_C2_2 * (2.f * zz - xx - yy);
    o[n++] = SH_C2_3 * xz; o[n++] = SH_C2_4 * (xx - yy);
    if (deg < 3) return n;
    o[n++] = SH_C3_0 * y * (3.f * xx - yy); o[n++] = SH_C3_1 * xy * z;
    o[n++] = SH_C3_2 * y * (4.f * zz - xx - yy); o[n++] = SH_C3_3 * z * (2.f * zz - 3.f * xx - 3.f * yy);
    o[n++] = SH_C3_4 * x * (4.f * zz - xx - yy); o[n++] = SH_C3_5 * z * (xx - yy);
    o[n++] = SH_C3_6 * x * (xx - 3.f * yy);
    return n;
}

// ── VJPs ─────────────────────────────────────────────────────────────────────

// pixel-position cotangent back to the world point, through the perspective divide
__device__ inline void pixel_vjp(const float *pm, const float *p, uint32_t W, uint32_t H, float gx, float gy,
                                 float *o) {
    float hx = pm[0] * p[0] + pm[1] * p[1] + pm[2] * p[2] + pm[3];
    float hy = pm[4] * p[0] + pm[5] * p[1] + pm[6] * p[2] + pm[7];
    float hw = pm[12] * p[0] + pm[13] * p[1] + pm[14] * p[2] + pm[15];
    float rw = 1.f / (hw + 1e-6f);
    float nx = 0.5f * (float)W * gx, ny = 0.5f * (float)H * gy;
    float q0 = nx * rw, q1 = ny * rw, q2 = 0.f, q3 = -(nx * hx + ny * hy) * rw * rw;
    o[0] = pm[0] * q0 + pm[4] * q1 + pm[8] * q2 + pm[12] * q3;
    o[1] = pm[1] * q0 + pm[5] * q1 + pm[9] * q2 + pm[13] * q3;
    o[2] = pm[2] * q0 + pm[6] * q1 + pm[10] * q2 + pm[14] * q3;
}

// d(conic)/d(cov2d) = -X G X, symmetric cotangent
__device__ inline void conic_vjp(const float *cn, const float *vc, float *o) {
    float X[2][2] = {{cn[0], cn[1]}, {cn[1], cn[2]}};
    float G[2][2] = {{vc[0], vc[1]}, {vc[1], vc[2]}};
    float P[2][2], Q[2][2];
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
        for (int i = 0; i < 2; i++) P[j][i] = X[0][i] * G[j][0] + X[1][i] * G[j][1];
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
        for (int i = 0; i < 2; i++) Q[j][i] = P[0][i] * X[j][0] + P[1][i] * X[j][1];
    o[0] = -1.f * Q[0][0];
    o[1] = -1.f * Q[1][0] + -1.f * Q[0][1];
    o[2] = -1.f * Q[1][1];
}

__device__ inline void ewa_vjp(const float *cov, const float *vm, float fx, float fy, float tfx, float tfy,
                               const float *vc2, const float *pv_in, float *v_mean, float *v_cov) {
    float pv[3] = {pv_in[0], pv_in[1], pv_in[2]};
    fov_clamp(pv, tfx, tfy);
    float rz = 1.f / pv[2], rz2 = rz * rz;
    m3 Wm = cols(vm[0], vm[4], vm[8], vm[1], vm[5], vm[9], vm[2], vm[6], vm[10]);
    m3 J = cols(fx * rz, 0.f, 0.f, 0.f, fy * rz, 0.f, -fx * pv[0] * rz2, -fy * pv[1] * rz2, 0.f);
    m3 V = cols(cov[0], cov[1], cov[2], cov[1], cov[3], cov[4], cov[2], cov[4], cov[5]);
    m3 G = cols(vc2[0], 0.5f * vc2[1], 0.f, 0.5f * vc2[1], vc2[2], 0.f, 0.f, 0.f, 0.f);
    m3 Tm = mm(J, Wm);
    m3 vV = mm(mm(tr(Tm), G), Tm);
    m3 A = mm(mm(G, Tm), tr(V)), B = mm(mm(tr(G), Tm), V);
    m3 vT;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) vT.a[c][r] = A.a[c][r] + B.a[c][r];
    v_cov[0] = vV.a[0][0];
    v_cov[1] = vV.a[0][1] + vV.a[1][0];
    v_cov[2] = vV.a[0][2] + vV.a[2][0];
    v_cov[3] = vV.a[1][1];
    v_cov[4] = vV.a[1][2] + vV.a[2][1];
    v_cov[5] = vV.a[2][2];
    m3 vJ = mm(vT, tr(Wm));
    float fr2 = fx * rz2, gr2 = fy * rz2, rz3 = rz2 * rz;
    float vt[3] = {-fr2 * vJ.a[2][0], -gr2 * vJ.a[2][1],
                   -fr2 * vJ.a[0][0] + 2.f * fx * pv[0] * rz3 * vJ.a[2][0] - gr2 * vJ.a[1][1] +
                       2.f * fy * pv[1] * rz3 * vJ.a[2][1]};
    v_mean[0] += dot3(vt, Wm.a[0]);
    v_mean[1] += dot3(vt, Wm.a[1]);
    v_mean[2] += dot3(vt, Wm.a[2]);
}

__device__ inline void quat_vjp(const float *q, const m3 &vR, float *o) {
    float s = 1.f / sqrtf(q[3] * q[3] + q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    float w = q[0] * s, x = q[1] * s, y = q[2] * s, z = q[3] * s;
#define R_(c, i) vR.a[c][i]
    float gw = 2.f * (x * (R_(1, 2) - R_(2, 1)) + y * (R_(2, 0) - R_(0, 2)) + z * (R_(0, 1) - R_(1, 0)));
    float gx = 2.f * (-2.f * x * (R_(1, 1) + R_(2, 2)) + y * (R_(0, 1) + R_(1, 0)) + z * (R_(0, 2) + R_(2, 0)) +
                      w * (R_(1, 2) - R_(2, 1)));
    float gy = 2.f * (x * (R_(0, 1) + R_(1, 0)) - 2.f * y * (R_(0, 0) + R_(2, 2)) + z * (R_(1, 2) + R_(2, 1)) +
                      w * (R_(2, 0) - R_(0, 2)));
    float gz = 2.f * (x * (R_(0, 2) + R_(2, 0)) + y * (R_(1, 2) + R_(2, 1)) - 2.f * z * (R_(0, 0) + R_(1, 1)) +
                      w * (R_(0, 1) - R_(1, 0)));
#undef R_
    float radial = gw * w + gx * x + gy * y + gz * z;
    o[0] = (gw - radial * w) * s;
    o[1] = (gx - radial * x) * s;
    o[2] = (gy - radial * y) * s;
    o[3] = (gz - radial * z) * s;
}

__device__ inline void cov3d_vjp(const float *scale, float glob, const float *q, const float *vc, float *v_scale,
                                 float *v_quat) {
    m3 vV = cols(vc[0], 0.5f * vc[1], 0.5f * vc[2], 0.5f * vc[1], vc[3], 0.5f * vc[4], 0.5f * vc[2], 0.5f * vc[4],
                 vc[5]);
    m3 R = rotation(q), S = diag(scale, glob);
    m3 M = mm(R, S);
    m3 vM = mm(vV, M);
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) vM.a[c][r] = vM.a[c][r] * 2.f;
    v_scale[0] = dot3(R.a[0], vM.a[0]);
    v_scale[1] = dot3(R.a[1], vM.a[1]);
    v_scale[2] = dot3(R.a[2], vM.a[2]);
    quat_vjp(q, mm(vM, S), v_quat);
}

}  // namespace msplat
