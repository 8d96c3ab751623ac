// msplat oracle — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of msplat's CPU backend (core/cpu/msplat_cpu.cpp, v1.1.4), written
// for this project as the parity checker of the HIP kernels. Only tests/, the smoke
// check and bench.py's cpu_baseline leg may load it; the product never does.
//
// Every function names the reference function it restates (file:line). The float
// instantiation keeps the reference's floating-point operation order (left-to-right
// sums, the same products), so that compiled with -ffp-contract=off it reproduces the
// reference bit for bit; the double instantiation is the FP64 renderer against which
// the FP32 outputs' errors are measured. The whole file is built with -ffp-contract=off.
//
// Pinning (what checks each function without trusting this file):
//   project/bin/raster/loss/backward/adam/stats/densify/knn/init/metrics/pack: the
//     golden vectors written by the compiled reference (oracle/_ref) in tests/golden,
//     compared bit-exactly for float;
//   projection/EWA/SH: closed forms in tests (rotation orthonormality, isotropic
//     Gaussian footprint, SH band-0 color);
//   compositing: closed form on one- and two-Gaussian pixels;
//   loss and every gradient: central finite differences of the double renderer;
//   radix sort: numpy stable argsort; kNN: brute force; PCG: integer known answers.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace orc {

constexpr int TILE = 16;  // binning tile edge, msplat_cpu.cpp:27

// Real SH basis constants, msplat_cpu.cpp:29-38 (values of the real spherical harmonics)
constexpr double kC0 = 0.28209479177387814;
constexpr double kC1 = 0.4886025119029199;
constexpr double kC2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792,
                           0.5462742152960396};
constexpr double kC3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
                           -0.4570457994644658, 1.445305721320277, -0.5900435899266435};
constexpr double kC4[9] = {2.5033429417967046, -1.7701307697799304, 0.9461746957575601, -0.6690465435572892,
                           0.10578554691520431, -0.6690465435572892, 0.47308734787878004, -1.7701307697799304,
                           0.6258357354491761};
// 11-tap Gaussian window, sigma 1.5 (msplat_cpu.cpp:40-45)
constexpr double kGauss[11] = {0.0010283801, 0.0075987581, 0.0360007721, 0.1093606895, 0.2130055377, 0.2660117249,
                               0.2130055377, 0.1093606895, 0.0360007721, 0.0075987581, 0.0010283801};

template <typename T> struct K {
    static T c(double v) { return (T)v; }
};

// ── small vector algebra (column-major 3x3, as the reference's MSL semantics) ─
template <typename T> struct v3 {
    T x, y, z;
};
template <typename T> inline v3<T> add(v3<T> a, v3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> inline v3<T> mul(v3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> inline T dot(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// a * s + c, unfused (the reference's fma3 is a plain multiply-add)
template <typename T> inline v3<T> madd(v3<T> a, T s, v3<T> c) { return {a.x * s + c.x, a.y * s + c.y, a.z * s + c.z}; }

template <typename T> struct m3 {
    v3<T> col[3];
    T &at(int c, int r) { return (&col[c].x)[r]; }
    T at(int c, int r) const { return (&col[c].x)[r]; }
};
template <typename T> inline m3<T> mm(const m3<T> &A, const m3<T> &B) {
    m3<T> R;
    for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++)
            R.at(j, i) = A.at(0, i) * B.at(j, 0) + A.at(1, i) * B.at(j, 1) + A.at(2, i) * B.at(j, 2);
    return R;
}
template <typename T> inline m3<T> tr(const m3<T> &A) {
    m3<T> R;
    for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) R.at(j, i) = A.at(i, j);
    return R;
}
template <typename T> inline m3<T> cols(T a0, T a1, T a2, T b0, T b1, T b2, T c0, T c1, T c2) {
    m3<T> R;
    R.col[0] = {a0, a1, a2};
    R.col[1] = {b0, b1, b2};
    R.col[2] = {c0, c1, c2};
    return R;
}

template <typename T> inline T sigmoid(T v) { return T(1) / (T(1) + std::exp(-v)); }

inline uint32_t sh_bases(uint32_t degree) {
    return degree == 0 ? 1 : degree == 1 ? 4 : degree == 2 ? 9 : degree == 3 ? 16 : 25;
}

// ── projection helpers (msplat_cpu.cpp:63-157) ────────────────────────────────

// Rotation of the stored quaternion (w, x, y, z), normalized; msplat_cpu.cpp:77-84
template <typename T> m3<T> rotation(const T *q) {
    T s = T(1) / std::sqrt(q[3] * q[3] + q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    T w = q[0] * s, x = q[1] * s, y = q[2] * s, z = q[3] * s;
    const T one = 1, two = 2;
    return cols<T>(one - two * (y * y + z * z), two * (x * y + w * z), two * (x * z - w * y),
                   two * (x * y - w * z), one - two * (x * x + z * z), two * (y * z + w * x),
                   two * (x * z + w * y), two * (y * z - w * x), one - two * (x * x + y * y));
}

template <typename T> m3<T> diag(v3<T> s, T g) {
    return cols<T>(g * s.x, 0, 0, 0, g * s.y, 0, 0, 0, g * s.z);
}

// Sigma = (R S)(R S)^T upper triangle; msplat_cpu.cpp:94-100
template <typename T> void covariance3d(v3<T> scale, T glob, const T *q, T *cov) {
    m3<T> M = mm(rotation(q), diag(scale, glob));
    m3<T> C = mm(M, tr(M));
    cov[0] = C.at(0, 0); cov[1] = C.at(0, 1); cov[2] = C.at(0, 2);
    cov[3] = C.at(1, 1); cov[4] = C.at(1, 2); cov[5] = C.at(2, 2);
}

// Camera-space point clamped to 1.3x the field of view (EWA Jacobian input)
template <typename T> v3<T> fov_clamp(v3<T> p, T tfx, T tfy) {
    T lx = K<T>::c(1.3) * tfx, ly = K<T>::c(1.3) * tfy;
    p.x = p.z * std::min(lx, std::max(-lx, p.x / p.z));
    p.y = p.z * std::min(ly, std::max(-ly, p.y / p.z));
    return p;
}

// EWA splat of the 3D covariance plus the 0.3 px low-pass; msplat_cpu.cpp:103-125
template <typename T> v3<T> ewa_cov2d(const T *cov, const T *vm, T fx, T fy, T tfx, T tfy, v3<T> pv) {
    pv = fov_clamp(pv, tfx, tfy);
    T rz = T(1) / pv.z, rz2 = rz * rz;
    T j00 = fx * rz, j11 = fy * rz, j20 = -fx * pv.x * rz2, j21 = -fy * pv.y * rz2;
    v3<T> r0{vm[0], vm[1], vm[2]}, r1{vm[4], vm[5], vm[6]}, r2{vm[8], vm[9], vm[10]};
    v3<T> t0 = add(mul(r0, j00), mul(r2, j20));
    v3<T> t1 = add(mul(r1, j11), mul(r2, j21));
    T a = cov[0], b = cov[1], c = cov[2], d = cov[3], e = cov[4], f = cov[5];
    v3<T> u0{t0.x * a + t0.y * b + t0.z * c, t0.x * b + t0.y * d + t0.z * e, t0.x * c + t0.y * e + t0.z * f};
    v3<T> u1{t1.x * a + t1.y * b + t1.z * c, t1.x * b + t1.y * d + t1.z * e, t1.x * c + t1.y * e + t1.z * f};
    const T lp = K<T>::c(0.3);
    return {dot(u0, t0) + lp, dot(u0, t1), dot(u1, t1) + lp};
}

// Conic (inverse 2D covariance) and 3-sigma radius; msplat_cpu.cpp:127-138
template <typename T> bool conic_radius(v3<T> c2, v3<T> &conic, T &radius) {
    T det = c2.x * c2.z - c2.y * c2.y;
    if (det == T(0)) return false;
    T inv = T(1) / det;
    conic = {c2.z * inv, -c2.y * inv, c2.x * inv};
    T mid = K<T>::c(0.5) * (c2.x + c2.z);
    T disc = std::sqrt(std::max(K<T>::c(0.1), mid * mid - det));
    radius = std::ceil(T(3) * std::sqrt(mid + disc));
    return true;
}

// Pixel position through the full projection, msplat_cpu.cpp:140-146
template <typename T> void to_pixel(const T *pm, v3<T> p, uint32_t W, uint32_t H, T cx, T cy, T &ox, T &oy) {
    T hx = pm[0] * p.x + pm[1] * p.y + pm[2] * p.z + pm[3];
    T hy = pm[4] * p.x + pm[5] * p.y + pm[6] * p.z + pm[7];
    T hw = pm[12] * p.x + pm[13] * p.y + pm[14] * p.z + pm[15];
    T rw = T(1) / (hw + K<T>::c(1e-6));
    const T half = K<T>::c(0.5);
    ox = half * (T)(int)W * (hx * rw) + cx - half;
    oy = half * (T)(int)H * (hy * rw) + cy - half;
}

// Inclusive-min / exclusive-max tile rectangle; msplat_cpu.cpp:148-157
template <typename T> void tile_rect(T cx, T cy, T rx, T ry, int tx, int ty, int &x0, int &y0, int &x1, int &y1) {
    T ccx = cx / (T)TILE, ccy = cy / (T)TILE, rrx = rx / (T)TILE, rry = ry / (T)TILE;
    x0 = std::min(std::max(0, (int)(ccx - rrx)), tx);
    x1 = std::min(std::max(0, (int)(ccx + rrx + T(1))), tx);
    y0 = std::min(std::max(0, (int)(ccy - rry)), ty);
    y1 = std::min(std::max(0, (int)(ccy + rry + T(1))), ty);
}

// SH basis values of band l >= 1 for one direction, in the reference's storage order
// and with its product grouping (msplat_cpu.cpp:160-190). out[b] multiplies rest[b].
template <typename T> int sh_basis(uint32_t deg, v3<T> d, T *out) {
    const T x = d.x, y = d.y, z = d.z;
    const T xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    auto C = [](double v) { return (T)v; };
    int n = 0;
    if (deg < 1) return n;
    // band 1 is evaluated as C1 * (-y r0 + z r1 - x r2) — see color_of; its basis
    // values are used only by the backward (msplat_cpu.cpp:1128)
    out[n++] = -C(kC1) * y; out[n++] = C(kC1) * z; out[n++] = -C(kC1) * x;
    if (deg < 2) return n;
    out[n++] = C(kC2[0]) * xy;
    out[n++] = C(kC2[1]) * yz;
    out[n++] = C(kC2[2]) * (T(2) * zz - xx - yy);
    out[n++] = C(kC2[3]) * xz;
    out[n++] = C(kC2[4]) * (xx - yy);
    if (deg < 3) return n;
    out[n++] = C(kC3[0]) * y * (T(3) * xx - yy);
    out[n++] = C(kC3[1]) * xy * z;
    out[n++] = C(kC3[2]) * y * (T(4) * zz - xx - yy);
    out[n++] = C(kC3[3]) * z * (T(2) * zz - T(3) * xx - T(3) * yy);
    out[n++] = C(kC3[4]) * x * (T(4) * zz - xx - yy);
    out[n++] = C(kC3[5]) * z * (xx - yy);
    out[n++] = C(kC3[6]) * x * (xx - T(3) * yy);
    if (deg < 4) return n;
    out[n++] = C(kC4[0]) * xy * (xx - yy);
    out[n++] = C(kC4[1]) * yz * (T(3) * xx - yy);
    out[n++] = C(kC4[2]) * xy * (T(7) * zz - T(1));
    out[n++] = C(kC4[3]) * yz * (T(7) * zz - T(3));
    out[n++] = C(kC4[4]) * (zz * (T(35) * zz - T(30)) + T(3));
    out[n++] = C(kC4[5]) * xz * (T(7) * zz - T(3));
    out[n++] = C(kC4[6]) * (xx - yy) * (T(7) * zz - T(1));
    out[n++] = C(kC4[7]) * xz * (xx - T(3) * yy);
    out[n++] = C(kC4[8]) * (xx * (xx - T(3) * yy) - yy * (T(3) * xx - yy));
    return n;
}

// Raw view-dependent color (no +0.5), msplat_cpu.cpp:160-190. Band sums are formed
// inside parentheses first and then added, as in the reference.
template <typename T> void color_of(uint32_t deg, v3<T> d, const T *dc, const T *rest, T *col) {
    for (int c = 0; c < 3; c++) col[c] = (T)kC0 * dc[c];
    if (deg < 1) return;
    const T x = d.x, y = d.y, z = d.z;
    const T xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    auto C = [](double v) { return (T)v; };
    for (int c = 0; c < 3; c++) {
        const T *r = rest + c;  // r[3*b] is basis b of channel c
        col[c] += C(kC1) * (-y * r[0] + z * r[3] - x * r[6]);
        if (deg < 2) continue;
        col[c] += (C(kC2[0]) * xy * r[9] + C(kC2[1]) * yz * r[12] + C(kC2[2]) * (T(2) * zz - xx - yy) * r[15] +
                   C(kC2[3]) * xz * r[18] + C(kC2[4]) * (xx - yy) * r[21]);
        if (deg < 3) continue;
        col[c] += (C(kC3[0]) * y * (T(3) * xx - yy) * r[24] + C(kC3[1]) * xy * z * r[27] +
                   C(kC3[2]) * y * (T(4) * zz - xx - yy) * r[30] +
                   C(kC3[3]) * z * (T(2) * zz - T(3) * xx - T(3) * yy) * r[33] +
                   C(kC3[4]) * x * (T(4) * zz - xx - yy) * r[36] + C(kC3[5]) * z * (xx - yy) * r[39] +
                   C(kC3[6]) * x * (xx - T(3) * yy) * r[42]);
        if (deg < 4) continue;
        col[c] += (C(kC4[0]) * xy * (xx - yy) * r[45] + C(kC4[1]) * yz * (T(3) * xx - yy) * r[48] +
                   C(kC4[2]) * xy * (T(7) * zz - T(1)) * r[51] + C(kC4[3]) * yz * (T(7) * zz - T(3)) * r[54] +
                   C(kC4[4]) * (zz * (T(35) * zz - T(30)) + T(3)) * r[57] +
                   C(kC4[5]) * xz * (T(7) * zz - T(3)) * r[60] +
                   C(kC4[6]) * (xx - yy) * (T(7) * zz - T(1)) * r[63] +
                   C(kC4[7]) * xz * (xx - T(3) * yy) * r[66] +
                   C(kC4[8]) * (xx * (xx - T(3) * yy) - yy * (T(3) * xx - yy)) * r[69]);
    }
}

template <typename T> inline v3<T> cam_space(const T *vm, const T *p) {
    return {vm[0] * p[0] + vm[1] * p[1] + vm[2] * p[2] + vm[3], vm[4] * p[0] + vm[5] * p[1] + vm[6] * p[2] + vm[7],
            vm[8] * p[0] + vm[9] * p[1] + vm[10] * p[2] + vm[11]};
}

template <typename T> inline v3<T> view_direction(const T *p, const T *cam) {
    v3<T> d{p[0] - cam[0], p[1] - cam[1], p[2] - cam[2]};
    return mul(d, T(1) / std::sqrt(dot(d, d)));
}

// ── 1. projection + SH (project_and_sh, msplat_cpu.cpp:452-500) ───────────────
template <typename T>
void project(int N, const T *means, const T *scales, T glob, const T *quats, const T *vm, const T *pm, T fx, T fy,
             T cx, T cy, uint32_t W, uint32_t H, int tiles_x, int tiles_y, T clip, uint32_t degree, uint32_t dtu,
             const T *cam, const T *dc, const T *rest, T *xys, T *depths, T *conics, T *colors, T *aabb,
             int32_t *radii) {
    const uint32_t nb = sh_bases(degree);
#pragma omp parallel for schedule(static)
    for (int g = 0; g < N; g++) {
        radii[g] = 0;
        const T *p = means + 3 * (size_t)g;
        v3<T> pv = cam_space(vm, p);
        if (pv.z <= clip) continue;
        v3<T> s{std::exp(scales[3 * g]), std::exp(scales[3 * g + 1]), std::exp(scales[3 * g + 2])};
        T cov[6];
        covariance3d(s, glob, quats + 4 * (size_t)g, cov);
        T tfx = K<T>::c(0.5) * (T)W / fx, tfy = K<T>::c(0.5) * (T)H / fy;
        v3<T> c2 = ewa_cov2d(cov, vm, fx, fy, tfx, tfy, pv);
        v3<T> conic;
        T radius;
        if (!conic_radius(c2, conic, radius)) continue;
        conics[3 * g] = conic.x; conics[3 * g + 1] = conic.y; conics[3 * g + 2] = conic.z;
        T px, py;
        to_pixel(pm, v3<T>{p[0], p[1], p[2]}, W, H, cx, cy, px, py);
        T ax = std::ceil(T(3) * std::sqrt(c2.x)), ay = std::ceil(T(3) * std::sqrt(c2.z));
        int x0, y0, x1, y1;
        tile_rect(px, py, ax, ay, tiles_x, tiles_y, x0, y0, x1, y1);
        if ((x1 - x0) * (y1 - y0) <= 0) continue;
        depths[g] = pv.z;
        radii[g] = (int)radius;
        xys[2 * g] = px; xys[2 * g + 1] = py;
        aabb[2 * g] = ax; aabb[2 * g + 1] = ay;
        color_of(dtu, view_direction(p, cam), dc + 3 * (size_t)g, rest + (size_t)(nb - 1) * 3 * g, colors + 3 * (size_t)g);
    }
}

// ── 2. binning (bin_and_sort, msplat_cpu.cpp:505-567) ─────────────────────────
template <typename T> inline uint64_t depth_key(T depth, int g) {
    float f = (float)depth;  // the key holds the fp32 bits even for the double renderer
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return ((uint64_t)b << 32) | (uint32_t)g;
}

template <typename T>
int64_t count_isects(int N, const T *xys, const T *aabb, const int32_t *radii, int tx, int ty, int32_t *tile_count) {
    int64_t total = 0;
    if (tile_count) std::fill(tile_count, tile_count + (size_t)tx * ty, 0);
    for (int g = 0; g < N; g++) {
        if (radii[g] <= 0) continue;
        int x0, y0, x1, y1;
        tile_rect(xys[2 * g], xys[2 * g + 1], aabb[2 * g], aabb[2 * g + 1], tx, ty, x0, y0, x1, y1);
        for (int y = y0; y < y1; y++)
            for (int x = x0; x < x1; x++) {
                if (tile_count) tile_count[y * tx + x]++;
                total++;
            }
    }
    return total;
}

// Per-tile lists sorted front to back by (depth bits, gaussian index), never truncated,
// and packed (x, y, sigmoid(opacity)), conic, raw color per intersection.
template <typename T>
void bin_sort(int N, const T *xys, const T *depths, const T *aabb, const int32_t *radii, const T *opac,
              const T *conics, const T *colors, int tx, int ty, int32_t *tile_start, int32_t *tile_end,
              int32_t *isect_gid, T *pk_xyo, T *pk_conic, T *pk_rgb, uint64_t *keys_out) {
    const int nt = tx * ty;
    std::vector<int32_t> cnt(nt);
    count_isects(N, xys, aabb, radii, tx, ty, cnt.data());
    std::vector<int64_t> off(nt + 1, 0);
    for (int t = 0; t < nt; t++) off[t + 1] = off[t] + cnt[t];
    if (off[nt] > std::numeric_limits<int32_t>::max()) throw std::runtime_error("more than 2^31 intersections");
    std::vector<uint64_t> bins((size_t)std::max<int64_t>(off[nt], 1));
    std::vector<int64_t> cur(off.begin(), off.end() - 1);
    for (int g = 0; g < N; g++) {
        if (radii[g] <= 0) continue;
        int x0, y0, x1, y1;
        tile_rect(xys[2 * g], xys[2 * g + 1], aabb[2 * g], aabb[2 * g + 1], tx, ty, x0, y0, x1, y1);
        uint64_t key = depth_key(depths[g], g);
        for (int y = y0; y < y1; y++)
            for (int x = x0; x < x1; x++) bins[cur[y * tx + x]++] = key;
    }
#pragma omp parallel for schedule(dynamic, 8)
    for (int t = 0; t < nt; t++) std::sort(bins.begin() + off[t], bins.begin() + off[t + 1]);
    for (int t = 0; t < nt; t++) {
        tile_start[t] = (int32_t)off[t];
        tile_end[t] = (int32_t)off[t + 1];
    }
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < off[nt]; k++) {
        int g = (int)(bins[k] & 0xFFFFFFFFull);
        isect_gid[k] = g;
        if (keys_out) keys_out[k] = bins[k];
        pk_xyo[3 * k] = xys[2 * g];
        pk_xyo[3 * k + 1] = xys[2 * g + 1];
        pk_xyo[3 * k + 2] = sigmoid(opac[g]);
        for (int c = 0; c < 3; c++) {
            pk_conic[3 * k + c] = conics[3 * (size_t)g + c];
            pk_rgb[3 * k + c] = colors[3 * (size_t)g + c];
        }
    }
}

// Gaussian exponent sigma = 1/2 (a dx^2 + c dy^2) + b dx dy, msplat_cpu.cpp:569-571
template <typename T> inline T splat_sigma(const T *conic, T dx, T dy) {
    return K<T>::c(0.5) * (conic[0] * dx * dx + conic[2] * dy * dy) + conic[1] * dx * dy;
}

template <typename T> struct Thr {
    static T sig_max() { return (T)5.55; }
    static T a_max() { return (T)0.999; }
    static T a_min() { return T(1) / T(255); }
    static T t_min() { return (T)1e-4; }
};
template <> struct Thr<float> {
    static float sig_max() { return 5.55f; }
    static float a_max() { return 0.999f; }
    static float a_min() { return 1.f / 255.f; }
    static float t_min() { return 1e-4f; }
};

// ── 3. forward compositing (rasterize_forward, msplat_cpu.cpp:574-626) ────────
// With `aux` (the prior-guided forward, rasterize_forward_aux): the expected-depth sum
// sum_i z_i vis_i (z = the gaussian's view depth, depths[isect_gid[k]]) and a per-pixel
// background (bg_of: the learned sky's colour, else the constant) written to bg_img.
template <typename T> struct AuxFwd {
    const T *depths = nullptr;
    const int32_t *gid = nullptr;
    T *out_depth = nullptr, *bg_img = nullptr;
    const T *sky_bg = nullptr;  // per-pixel background (W x H x 3), or null: the constant
};

template <typename T>
void raster_forward(uint32_t W, uint32_t H, int tx, int ty, const int32_t *ts, const int32_t *te, const T *pk_xyo,
                    const T *pk_conic, const T *pk_rgb, const T *bg, T *out, T *final_T, int32_t *final_idx,
                    const AuxFwd<T> *aux = nullptr) {
    const int nt = tx * ty;
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < nt; t++) {
        const int bx = t % tx, by = t / tx;
        for (int i = by * TILE; i < std::min((int)H, (by + 1) * TILE); i++)
            for (int j = bx * TILE; j < std::min((int)W, (bx + 1) * TILE); j++) {
                const T px = (T)j, py = (T)i;
                T Tr = 1, r = 0, gg = 0, b = 0, depth = 0;
                int last = ts[t] - 1;
                for (int k = ts[t]; k < te[t]; k++) {
                    const T *xyo = pk_xyo + 3 * (size_t)k;
                    T s = splat_sigma(pk_conic + 3 * (size_t)k, xyo[0] - px, xyo[1] - py);
                    if (s < T(0) || s >= Thr<T>::sig_max()) continue;
                    T a = std::min(Thr<T>::a_max(), xyo[2] * std::exp(-s));
                    if (a < Thr<T>::a_min()) continue;
                    T nT = Tr * (T(1) - a);
                    if (nT <= Thr<T>::t_min()) {
                        last = k - 1;
                        break;
                    }
                    T vis = a * Tr;
                    const T *c = pk_rgb + 3 * (size_t)k;
                    r = std::max(c[0] + K<T>::c(0.5), T(0)) * vis + r;
                    gg = std::max(c[1] + K<T>::c(0.5), T(0)) * vis + gg;
                    b = std::max(c[2] + K<T>::c(0.5), T(0)) * vis + b;
                    if (aux) depth = depth + aux->depths[aux->gid[k]] * vis;
                    Tr = nT;
                    last = k;
                }
                const size_t pid = (size_t)i * W + j;
                final_T[pid] = Tr;
                final_idx[pid] = last;
                const T *bgp = aux && aux->sky_bg ? aux->sky_bg + 3 * pid : bg;
                if (aux) {
                    aux->out_depth[pid] = depth;
                    for (int c = 0; c < 3; c++) aux->bg_img[3 * pid + c] = bgp[c];
                }
                out[3 * pid] = std::min(std::max(bgp[0] * Tr + r, T(0)), T(1));
                out[3 * pid + 1] = std::min(std::max(bgp[1] * Tr + gg, T(0)), T(1));
                out[3 * pid + 2] = std::min(std::max(bgp[2] * Tr + b, T(0)), T(1));
            }
    }
}

// ── 4. L1 + SSIM loss and its image gradient (msplat_cpu.cpp:669-785) ─────────
// Zero-padded 11x11 SSIM between render y and target x; returns the summed per-pixel
// loss (per-pixel values in float, accumulated in double per row, rows in order).
double loss_fwd_bwd(uint32_t W, uint32_t H, const float *ren, const float *gt, float w, float inv_n, float *v_ren) {
    const int R = 5, WIN = 11;
    float G[11];
    for (int k = 0; k < 11; k++) G[k] = (float)kGauss[k];
    const float C1 = 0.0001f, C2 = 0.0009f;
    std::vector<float> mom((size_t)W * H * 15), der((size_t)W * H * 15);
    // horizontal moments mu_x, mu_y, E[x^2], E[y^2], E[xy] per channel (ssim_h_fwd)
#pragma omp parallel for schedule(static)
    for (int y = 0; y < (int)H; y++)
        for (int x = 0; x < (int)W; x++)
            for (int c = 0; c < 3; c++) {
                float m[5] = {0, 0, 0, 0, 0};
                for (int k = 0; k < WIN; k++) {
                    int sx = x - R + k;
                    float gv = 0.f, rv = 0.f;
                    if (sx >= 0 && sx < (int)W) {
                        size_t i = ((size_t)y * W + sx) * 3 + c;
                        gv = gt[i];
                        rv = ren[i];
                    }
                    m[0] += G[k] * gv; m[1] += G[k] * rv;
                    m[2] += G[k] * gv * gv; m[3] += G[k] * rv * rv; m[4] += G[k] * gv * rv;
                }
                std::memcpy(&mom[((size_t)y * W + x) * 15 + c * 5], m, sizeof m);
            }
    // vertical pass, SSIM, dSSIM/d(moment) fields, horizontal back-convolution
    std::vector<double> rows(H, 0.0);
#pragma omp parallel
    {
        const int ext = (int)W + 2 * R;
        std::vector<float> f1(ext), f2(ext), f3(ext), sp(W), lp(W);
#pragma omp for schedule(static)
        for (int y = 0; y < (int)H; y++) {
            std::fill(sp.begin(), sp.end(), 0.f);
            std::fill(lp.begin(), lp.end(), 0.f);
            for (int c = 0; c < 3; c++) {
                for (int e = 0; e < ext; e++) {
                    const int x = e - R;
                    if (x < 0 || x >= (int)W) {
                        f1[e] = f2[e] = f3[e] = 0.f;
                        continue;
                    }
                    float s[5] = {0, 0, 0, 0, 0};
                    for (int k = 0; k < WIN; k++) {
                        int sy = y - R + k;
                        if (sy < 0 || sy >= (int)H) continue;
                        const float *h = &mom[((size_t)sy * W + x) * 15 + c * 5];
                        for (int q = 0; q < 5; q++) s[q] += G[k] * h[q];
                    }
                    float mx = s[0], my = s[1];
                    float vx = s[2] - mx * mx, vy = s[3] - my * my, cxy = s[4] - mx * my;
                    float A = 2.0f * mx * my + C1, B = 2.0f * cxy + C2;
                    float Cd = mx * mx + my * my + C1, D = vx + vy + C2;
                    float iCD = 1.0f / (Cd * D);
                    float dmu = 2.0f * B * (mx * Cd - A * my) / (Cd * Cd * D);
                    float dvy = -A * B * iCD / D, dcv = 2.0f * A * iCD;
                    f1[e] = dmu - 2.0f * my * dvy - mx * dcv;
                    f2[e] = 2.0f * dvy;
                    f3[e] = dcv;
                    size_t i = ((size_t)y * W + x) * 3 + c;
                    sp[x] += (A * B) / (Cd * D);
                    lp[x] += std::fabs(gt[i] - ren[i]);
                }
                for (int x = 0; x < (int)W; x++) {
                    float h1 = 0, h2 = 0, h3 = 0;
                    for (int d = 0; d < WIN; d++) {
                        float g = G[WIN - 1 - d];
                        h1 += g * f1[x + d];
                        h2 += g * f2[x + d];
                        h3 += g * f3[x + d];
                    }
                    float *o = &der[((size_t)y * W + x) * 15 + c * 5];
                    o[0] = h1; o[1] = h2; o[2] = h3;
                }
            }
            double acc = 0.0;
            for (int x = 0; x < (int)W; x++) acc += w * (1.0f - sp[x] / 3.0f) + (1.0f - w) * lp[x] / 3.0f;
            rows[y] = acc;
        }
    }
    double total = 0.0;
    for (double r : rows) total += r;
    // vertical back-convolution and the image gradient (ssim_v_bwd)
#pragma omp parallel for schedule(static)
    for (int y = 0; y < (int)H; y++)
        for (int x = 0; x < (int)W; x++)
            for (int c = 0; c < 3; c++) {
                float a1 = 0, a2 = 0, a3 = 0;
                for (int d = 0; d < WIN; d++) {
                    int sy = y - R + d;
                    if (sy < 0 || sy >= (int)H) continue;
                    const float *o = &der[((size_t)sy * W + x) * 15 + c * 5];
                    float g = G[WIN - 1 - d];
                    a1 += g * o[0]; a2 += g * o[1]; a3 += g * o[2];
                }
                size_t i = ((size_t)y * W + x) * 3 + c;
                float rv = ren[i], gv = gt[i];
                float vs = a1 + rv * a2 + gv * a3;
                float vl = (gv > rv) ? -1.0f : ((gv < rv) ? 1.0f : 0.0f);
                v_ren[i] = inv_n * (-w * vs + (1.0f - w) * vl);
            }
    return total;
}

// ── 5. backward compositing (rasterize_backward, msplat_cpu.cpp:959-1031) ─────
// Per-intersection gradient slots, 10 floats: v_xy(2) v_conic(3) v_rgb(3) v_logit v_depth.
// With `aux` (rasterize_backward with the aux inputs, msplat_cpu.cpp:963-1025): the per
// pixel background of the forward, and the prior losses' gradients with respect to the
// expected-depth sum D and alpha A (v_aux: 2 per pixel), which add to v_alpha and give
// each intersection a depth gradient (slot 9).
template <typename T> struct AuxBwd {
    const T *depths = nullptr;
    const int32_t *gid = nullptr;
    const T *bg_img = nullptr, *v_aux = nullptr;
};

template <typename T>
void raster_backward(uint32_t W, uint32_t H, int tx, int ty, const int32_t *ts, const int32_t *te, const T *pk_xyo,
                     const T *pk_conic, const T *pk_rgb, const T *bg, const T *final_T, const int32_t *final_idx,
                     const T *v_out, T *slots, const AuxBwd<T> *aux = nullptr) {
    const int nt = tx * ty;
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < nt; t++) {
        const int s0 = ts[t], s1 = te[t];
        T *sl = slots + (size_t)s0 * 10;
        std::fill(sl, sl + (size_t)(s1 - s0) * 10, T(0));
        if (s1 == s0) continue;
        const int bx = t % tx, by = t / tx;
        for (int i = by * TILE; i < std::min((int)H, (by + 1) * TILE); i++)
            for (int j = bx * TILE; j < std::min((int)W, (bx + 1) * TILE); j++) {
                const size_t pid = (size_t)i * W + j;
                const T px = (T)j, py = (T)i;
                const T Tf = final_T[pid];
                T Tr = Tf;
                v3<T> buf{0, 0, 0};
                T buf_d = 0, buf_a = 0;
                const v3<T> vo{v_out[3 * pid], v_out[3 * pid + 1], v_out[3 * pid + 2]};
                const T *bgp = aux ? aux->bg_img + 3 * pid : bg;
                const T v_d = aux ? aux->v_aux[2 * pid] : T(0), v_ax = aux ? aux->v_aux[2 * pid + 1] : T(0);
                const v3<T> tbg{Tf * bgp[0], Tf * bgp[1], Tf * bgp[2]};
                for (int k = std::min(final_idx[pid], s1 - 1); k >= s0; k--) {
                    const T *xyo = pk_xyo + 3 * (size_t)k;
                    const T *cn = pk_conic + 3 * (size_t)k;
                    const T dx = xyo[0] - px, dy = xyo[1] - py;
                    T s = splat_sigma(cn, dx, dy);
                    if (s < T(0) || s >= Thr<T>::sig_max()) continue;
                    T a = std::min(Thr<T>::a_max(), xyo[2] * std::exp(-s));
                    if (a < Thr<T>::a_min()) continue;
                    T ra = T(1) / (T(1) - a);
                    Tr *= ra;
                    const T fac = a * Tr;
                    const T *raw = pk_rgb + 3 * (size_t)k;
                    const T h = K<T>::c(0.5);
                    v3<T> rgb{std::max(raw[0] + h, T(0)), std::max(raw[1] + h, T(0)), std::max(raw[2] + h, T(0))};
                    // <rgb T - (buf + T_final bg) / (1 - alpha), v_out>
                    v3<T> nb{-buf.x, -buf.y, -buf.z};
                    v3<T> e = madd(rgb, Tr, madd(nb, ra, mul(tbg, -ra)));
                    T v_a = dot(e, vo);
                    T z = 0;
                    if (aux) {
                        z = aux->depths[aux->gid[k]];
                        v_a += (z * Tr - buf_d * ra) * v_d + (Tr - buf_a * ra) * v_ax;
                    }
                    buf = madd(rgb, fac, buf);
                    T *g = sl + (size_t)(k - s0) * 10;
                    if (raw[0] + h >= T(0)) g[5] += fac * vo.x;
                    if (raw[1] + h >= T(0)) g[6] += fac * vo.y;
                    if (raw[2] + h >= T(0)) g[7] += fac * vo.z;
                    if (aux) {
                        buf_d += z * fac;
                        buf_a += fac;
                        g[9] += fac * v_d;
                    }
                    if (a < Thr<T>::a_max()) {
                        const T vs = -a * v_a;
                        g[2] += h * vs * dx * dx;
                        g[3] += h * vs * dx * dy;
                        g[4] += h * vs * dy * dy;
                        g[0] += vs * (cn[0] * dx + cn[1] * dy);
                        g[1] += vs * (cn[1] * dx + cn[2] * dy);
                        g[8] += -vs * (T(1) - xyo[2]);
                    }
                }
            }
    }
}

// Fixed-order per-gaussian sum of the slots (reduce_isect_grads, msplat_cpu.cpp:1034-1064)
template <typename T>
void reduce_slots(int N, int64_t M, const int32_t *isect_gid, const T *slots, T *v_xy, T *v_conic, T *v_rgb,
                  T *v_opac, T *v_depth) {
    std::vector<int64_t> off((size_t)N + 1, 0);
    for (int64_t k = 0; k < M; k++) off[isect_gid[k] + 1]++;
    for (int g = 0; g < N; g++) off[g + 1] += off[g];
    std::vector<int64_t> list((size_t)std::max<int64_t>(M, 1)), cur(off.begin(), off.end() - 1);
    for (int64_t k = 0; k < M; k++) list[cur[isect_gid[k]]++] = k;
#pragma omp parallel for schedule(static)
    for (int g = 0; g < N; g++) {
        T s[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t q = off[g]; q < off[g + 1]; q++)
            for (int c = 0; c < 10; c++) s[c] += slots[(size_t)list[q] * 10 + c];
        v_xy[2 * g] = s[0]; v_xy[2 * g + 1] = s[1];
        for (int c = 0; c < 3; c++) { v_conic[3 * g + c] = s[2 + c]; v_rgb[3 * g + c] = s[5 + c]; }
        v_opac[g] = s[8];
        v_depth[g] = s[9];
    }
}

// ── 6. projection VJPs (msplat_cpu.cpp:197-291) ───────────────────────────────
template <typename T> v3<T> pixel_vjp(const T *pm, const T *p, uint32_t W, uint32_t H, T gx, T gy) {
    T hx = pm[0] * p[0] + pm[1] * p[1] + pm[2] * p[2] + pm[3];
    T hy = pm[4] * p[0] + pm[5] * p[1] + pm[6] * p[2] + pm[7];
    T hw = pm[12] * p[0] + pm[13] * p[1] + pm[14] * p[2] + pm[15];
    (void)hw;
    T rw = T(1) / (hw + K<T>::c(1e-6));
    T nx = K<T>::c(0.5) * (T)W * gx, ny = K<T>::c(0.5) * (T)H * gy;
    T q0 = nx * rw, q1 = ny * rw, q2 = 0, q3 = -(nx * hx + ny * hy) * rw * rw;
    return {pm[0] * q0 + pm[4] * q1 + pm[8] * q2 + pm[12] * q3, pm[1] * q0 + pm[5] * q1 + pm[9] * q2 + pm[13] * q3,
            pm[2] * q0 + pm[6] * q1 + pm[10] * q2 + pm[14] * q3};
}

// d(conic)/d(cov2d): -X G X for X = conic, G = symmetric cotangent (msplat_cpu.cpp:207-214)
template <typename T> v3<T> conic_vjp(const T *cn, const T *vc) {
    // 2x2 column-major products, left to right
    T X[2][2] = {{cn[0], cn[1]}, {cn[1], cn[2]}};
    T G[2][2] = {{vc[0], vc[1]}, {vc[1], vc[2]}};
    T P[2][2], Q[2][2];
    for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) P[j][i] = X[0][i] * G[j][0] + X[1][i] * G[j][1];
    for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) Q[j][i] = P[0][i] * X[j][0] + P[1][i] * X[j][1];
    const T m1 = -1;
    return {m1 * Q[0][0], m1 * Q[1][0] + m1 * Q[0][1], m1 * Q[1][1]};
}

// EWA VJP: d(cov2d)/d(cov3d) and the mean term through the Jacobian (msplat_cpu.cpp:216-255)
template <typename T>
void ewa_vjp(const T *cov, const T *vm, T fx, T fy, T tfx, T tfy, v3<T> vc2, v3<T> pv, T *v_mean, T *v_cov) {
    pv = fov_clamp(pv, tfx, tfy);
    T rz = T(1) / pv.z, rz2 = rz * rz;
    m3<T> Wm = cols<T>(vm[0], vm[4], vm[8], vm[1], vm[5], vm[9], vm[2], vm[6], vm[10]);
    m3<T> J = cols<T>(fx * rz, 0, 0, 0, fy * rz, 0, -fx * pv.x * rz2, -fy * pv.y * rz2, 0);
    m3<T> V = cols<T>(cov[0], cov[1], cov[2], cov[1], cov[3], cov[4], cov[2], cov[4], cov[5]);
    const T h = K<T>::c(0.5);
    m3<T> G = cols<T>(vc2.x, h * vc2.y, 0, h * vc2.y, vc2.z, 0, 0, 0, 0);
    m3<T> Tm = mm(J, Wm);
    m3<T> vV = mm(mm(tr(Tm), G), Tm);
    m3<T> A = mm(mm(G, Tm), tr(V)), B = mm(mm(tr(G), Tm), V);
    m3<T> vT;
    for (int c = 0; c < 3; c++) vT.col[c] = add(A.col[c], B.col[c]);
    v_cov[0] = vV.at(0, 0);
    v_cov[1] = vV.at(0, 1) + vV.at(1, 0);
    v_cov[2] = vV.at(0, 2) + vV.at(2, 0);
    v_cov[3] = vV.at(1, 1);
    v_cov[4] = vV.at(1, 2) + vV.at(2, 1);
    v_cov[5] = vV.at(2, 2);
    m3<T> vJ = mm(vT, tr(Wm));
    T fr2 = fx * rz2, gr2 = fy * rz2, rz3 = rz2 * rz;
    v3<T> vt{-fr2 * vJ.at(2, 0), -gr2 * vJ.at(2, 1),
             -fr2 * vJ.at(0, 0) + T(2) * fx * pv.x * rz3 * vJ.at(2, 0) - gr2 * vJ.at(1, 1) +
                 T(2) * fy * pv.y * rz3 * vJ.at(2, 1)};
    v_mean[0] += dot(vt, Wm.col[0]);
    v_mean[1] += dot(vt, Wm.col[1]);
    v_mean[2] += dot(vt, Wm.col[2]);
}

// d/d(stored quaternion) through the normalization (msplat_cpu.cpp:259-274)
template <typename T> void quat_vjp(const T *q, const m3<T> &vR, T *out) {
    T s = T(1) / std::sqrt(q[3] * q[3] + q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    T w = q[0] * s, x = q[1] * s, y = q[2] * s, z = q[3] * s;
    auto r = [&](int c, int i) { return vR.at(c, i); };
    const T two = 2;
    T gw = two * (x * (r(1, 2) - r(2, 1)) + y * (r(2, 0) - r(0, 2)) + z * (r(0, 1) - r(1, 0)));
    T gx = two * (-two * x * (r(1, 1) + r(2, 2)) + y * (r(0, 1) + r(1, 0)) + z * (r(0, 2) + r(2, 0)) +
                  w * (r(1, 2) - r(2, 1)));
    T gy = two * (x * (r(0, 1) + r(1, 0)) - two * y * (r(0, 0) + r(2, 2)) + z * (r(1, 2) + r(2, 1)) +
                  w * (r(2, 0) - r(0, 2)));
    T gz = two * (x * (r(0, 2) + r(2, 0)) + y * (r(1, 2) + r(2, 1)) - two * z * (r(0, 0) + r(1, 1)) +
                  w * (r(0, 1) - r(1, 0)));
    T radial = gw * w + gx * x + gy * y + gz * z;
    out[0] = (gw - radial * w) * s;
    out[1] = (gx - radial * x) * s;
    out[2] = (gy - radial * y) * s;
    out[3] = (gz - radial * z) * s;
}

// d(cov3d)/d(scale, quat) (msplat_cpu.cpp:276-291)
template <typename T> void cov3d_vjp(v3<T> scale, T glob, const T *q, const T *vc, T *v_scale, T *v_quat) {
    const T h = K<T>::c(0.5);
    m3<T> vV = cols<T>(vc[0], h * vc[1], h * vc[2], h * vc[1], vc[3], h * vc[4], h * vc[2], h * vc[4], vc[5]);
    m3<T> R = rotation(q), S = diag(scale, glob);
    m3<T> M = mm(R, S);
    m3<T> vM = mm(vV, M);
    for (int c = 0; c < 3; c++) vM.col[c] = mul(vM.col[c], T(2));
    v_scale[0] = dot(R.col[0], vM.col[0]);
    v_scale[1] = dot(R.col[1], vM.col[1]);
    v_scale[2] = dot(R.col[2], vM.col[2]);
    quat_vjp(q, mm(vM, S), v_quat);
}

// Adam step in the reference's bc2_sqrt form (msplat_cpu.cpp:293-300)
template <typename T> inline void adam(T &p, T &m1, T &m2, T g, T step, T b1, T b2, T bc2, T eps) {
    T m = b1 * m1 + (T(1) - b1) * g;
    T v = b2 * m2 + (T(1) - b2) * g * g;
    p -= step * m / (std::sqrt(v) / bc2 + eps);
    m1 = m;
    m2 = v;
}

// Projection/SH backward with the SH Adam fused (project_and_sh_backward_adam,
// msplat_cpu.cpp:1068-1139). Writes v_mean3d, v_scale (log-space), v_quat; updates dc and
// rest (and their moments) in place for the visible gaussians only.
template <typename T>
void project_backward_adam(int N, const T *means, const T *scales, T glob, const T *quats, const T *vm, const T *pm,
                           T fx, T fy, uint32_t W, uint32_t H, uint32_t degree, uint32_t dtu, const T *cam,
                           const int32_t *radii, const T *conics, const T *v_xy, const T *v_conic, const T *v_rgb,
                           const T *v_depth, T *v_mean, T *v_scale, T *v_quat, T *dc, T *rest, T *dc_m1, T *dc_m2,
                           T *rest_m1, T *rest_m2, T dc_step, T dc_bc2, T rest_step, T rest_bc2, T b1, T b2, T eps) {
    const uint32_t nb = sh_bases(degree);
#pragma omp parallel for schedule(static)
    for (int g = 0; g < N; g++) {
        T *om = v_mean + 3 * (size_t)g, *os = v_scale + 3 * (size_t)g, *oq = v_quat + 4 * (size_t)g;
        if (radii[g] <= 0) {
            om[0] = om[1] = om[2] = 0;
            os[0] = os[1] = os[2] = 0;
            oq[0] = oq[1] = oq[2] = oq[3] = 0;
            continue;
        }
        const T *p = means + 3 * (size_t)g;
        v3<T> a = pixel_vjp(pm, p, W, H, v_xy[2 * g], v_xy[2 * g + 1]);
        a = add(a, mul(v3<T>{vm[8], vm[9], vm[10]}, v_depth[g]));
        om[0] = a.x; om[1] = a.y; om[2] = a.z;
        v3<T> vc2 = conic_vjp(conics + 3 * (size_t)g, v_conic + 3 * (size_t)g);
        v3<T> es{std::exp(scales[3 * g]), std::exp(scales[3 * g + 1]), std::exp(scales[3 * g + 2])};
        const T *q = quats + 4 * (size_t)g;
        T cov[6];
        covariance3d(es, glob, q, cov);
        T tfx = K<T>::c(0.5) * (T)W / fx, tfy = K<T>::c(0.5) * (T)H / fy;
        T vcov[6];
        ewa_vjp(cov, vm, fx, fy, tfx, tfy, vc2, cam_space(vm, p), om, vcov);
        cov3d_vjp(es, glob, q, vcov, os, oq);
        os[0] *= es.x; os[1] *= es.y; os[2] *= es.z;

        v3<T> d = view_direction(p, cam);
        const T gc[3] = {v_rgb[3 * g], v_rgb[3 * g + 1], v_rgb[3 * g + 2]};
        for (int c = 0; c < 3; c++) {
            size_t i = 3 * (size_t)g + c;
            adam(dc[i], dc_m1[i], dc_m2[i], (T)kC0 * gc[c], dc_step, b1, b2, dc_bc2, eps);
        }
        if (dtu < 1) continue;
        T basis[24];
        int nbasis = sh_basis(std::min<uint32_t>(dtu, 3), d, basis);
        const size_t base = (size_t)(nb - 1) * 3 * g;
        for (int b = 0; b < nbasis; b++)
            for (int c = 0; c < 3; c++) {
                size_t i = base + 3 * b + c;
                adam(rest[i], rest_m1[i], rest_m2[i], basis[b] * gc[c], rest_step, b1, b2, rest_bc2, eps);
            }
    }
}

// Densification statistics (msplat_cpu.cpp:1311-1324)
template <typename T>
void grad_stats(int N, const int32_t *radii, const T *v_xy, T inv_max_dim, T *vis, T *gn, T *ms) {
    for (int g = 0; g < N; g++) {
        if (radii[g] <= 0) continue;
        vis[g] += T(1);
        T gx = v_xy[2 * g], gy = v_xy[2 * g + 1];
        gn[g] += std::sqrt(gx * gx + gy * gy);
        ms[g] = std::max(ms[g], (T)radii[g] * inv_max_dim);
    }
}

// ── counter-based RNG (msplat_cpu.cpp:304-323): PCG hash (Jarzynski & Olano 2020),
//    24-bit uniforms in (0, 1), Box-Muller normals ──────────────────────────────
inline uint32_t pcg(uint32_t v) {
    uint32_t st = v * 747796405u + 2891336453u;
    uint32_t w = ((st >> ((st >> 28u) + 4u)) ^ st) * 277803737u;
    return (w >> 22u) ^ w;
}
inline float uniform01(uint32_t seed, uint32_t ctr) {
    uint32_t h = pcg(seed ^ pcg(ctr));
    return ((float)(h >> 8) + 0.5f) * (1.0f / 16777216.0f);
}
inline void normal3(uint32_t seed, uint32_t stream, float *out) {
    uint32_t c = stream * 4u;
    float u0 = uniform01(seed, c), u1 = uniform01(seed, c + 1u), u2 = uniform01(seed, c + 2u),
          u3 = uniform01(seed, c + 3u);
    float r0 = std::sqrt(-2.0f * std::log(u0)), r1 = std::sqrt(-2.0f * std::log(u2));
    const float tau = 6.28318530717958647692f;
    out[0] = r0 * std::cos(tau * u1);
    out[1] = r0 * std::sin(tau * u1);
    out[2] = r1 * std::cos(tau * u3);
}

// ── densify: classify, split, duplicate, cull, compact (msplat_cpu.cpp:1330-1465) ─
// bufs[0..5] = means, scales, quats, dc, rest, opacity; bufs[6..11] first moments,
// bufs[12..17] second moments, all with capacity >= 3N rows. Masks are returned when
// the out pointers are non-null: split/dup (N), keep (N_new).
int densify(int N, int cap, float grad_thr, float size_thr, float screen_thr, int check_screen, float cull_alpha,
            float cull_scale, float cull_screen, int check_huge, const float *gn, const float *vc, const float *ms,
            float half_max_dim, float **bufs, int fr_stride, uint32_t seed, uint8_t *split_out, uint8_t *dup_out,
            uint8_t *keep_out, int *n_new_out) {
    if (3 * N > cap) throw std::runtime_error("densify: 3*N exceeds capacity");
    const float shrink = std::log(1.6f);
    float *means = bufs[0], *scales = bufs[1], *quats = bufs[2], *fdc = bufs[3], *frest = bufs[4], *opac = bufs[5];
    std::vector<uint8_t> split(N), dup(N);
    auto max_exp_scale = [&](int i) {
        return std::max(std::max(std::exp(scales[3 * i]), std::exp(scales[3 * i + 1])), std::exp(scales[3 * i + 2]));
    };
    for (int i = 0; i < N; i++) {
        split[i] = dup[i] = 0;
        if (vc[i] <= 0.0f) continue;
        bool high = (gn[i] / vc[i]) * half_max_dim > grad_thr;
        bool large = max_exp_scale(i) > size_thr;
        bool sp = large || (check_screen && ms[i] > screen_thr);
        split[i] = (sp && high) ? 1 : 0;
        dup[i] = (!large && high) ? 1 : 0;
    }
    std::vector<int> sord(N, -1), dord(N, -1);
    int ns = 0, nd = 0;
    for (int i = 0; i < N; i++) {
        if (split[i]) sord[i] = ns++;
        if (dup[i]) dord[i] = nd++;
    }
    const int width[6] = {3, 3, 4, 3, fr_stride, 1};
    auto copy_row = [&](int dst, int src) {
        for (int b = 0; b < 6; b++)
            std::memcpy(bufs[b] + (size_t)dst * width[b], bufs[b] + (size_t)src * width[b], sizeof(float) * width[b]);
    };
    auto clear_moments = [&](int row) {
        for (int b = 6; b < 18; b++) std::fill(bufs[b] + (size_t)row * width[b % 6], bufs[b] + (size_t)(row + 1) * width[b % 6], 0.f);
    };
    for (int i = 0; i < N; i++) {
        if (!split[i]) continue;
        float qw = quats[4 * i], qx = quats[4 * i + 1], qy = quats[4 * i + 2], qz = quats[4 * i + 3];
        float ql = std::sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
        qw /= ql; qx /= ql; qy /= ql; qz /= ql;
        float sx = std::exp(scales[3 * i]), sy = std::exp(scales[3 * i + 1]), sz = std::exp(scales[3 * i + 2]);
        for (int k = 0; k < 2; k++) {
            const int ch = N + 2 * sord[i] + k;
            float n[3];
            normal3(seed, (uint32_t)(sord[i] * 2 + k), n);
            float r0 = n[0] * sx, r1 = n[1] * sy, r2 = n[2] * sz;
            float o0 = (1 - 2 * (qy * qy + qz * qz)) * r0 + 2 * (qx * qy - qw * qz) * r1 + 2 * (qx * qz + qw * qy) * r2;
            float o1 = 2 * (qx * qy + qw * qz) * r0 + (1 - 2 * (qx * qx + qz * qz)) * r1 + 2 * (qy * qz - qw * qx) * r2;
            float o2 = 2 * (qx * qz - qw * qy) * r0 + 2 * (qy * qz + qw * qx) * r1 + (1 - 2 * (qx * qx + qy * qy)) * r2;
            copy_row(ch, i);
            means[3 * ch] = means[3 * i] + o0;
            means[3 * ch + 1] = means[3 * i + 1] + o1;
            means[3 * ch + 2] = means[3 * i + 2] + o2;
            for (int j = 0; j < 3; j++) scales[3 * ch + j] = scales[3 * i + j] - shrink;
            clear_moments(ch);
        }
        for (int j = 0; j < 3; j++) scales[3 * i + j] -= shrink;
    }
    for (int i = 0; i < N; i++) {
        if (!dup[i]) continue;
        const int dst = N + 2 * ns + dord[i];
        copy_row(dst, i);
        clear_moments(dst);
    }
    const int Nn = N + 2 * ns + nd;
    std::vector<uint8_t> keep(Nn);
    for (int i = 0; i < Nn; i++) {
        float o = 1.0f / (1.0f + std::exp(-opac[i]));
        bool cull = o < cull_alpha;
        if (i < N && split[i]) cull = true;
        if (check_huge) {
            if (max_exp_scale(i) > cull_scale) cull = true;
            if (check_screen && i < N && ms[i] > cull_screen) cull = true;
        }
        keep[i] = cull ? 0 : 1;
    }
    int out = 0;
    for (int i = 0; i < Nn; i++) {
        if (!keep[i]) continue;
        if (out != i)
            for (int b = 0; b < 18; b++)
                std::memmove(bufs[b] + (size_t)out * width[b % 6], bufs[b] + (size_t)i * width[b % 6],
                             sizeof(float) * width[b % 6]);
        out++;
    }
    if (split_out) std::memcpy(split_out, split.data(), N);
    if (dup_out) std::memcpy(dup_out, dup.data(), N);
    if (keep_out) std::memcpy(keep_out, keep.data(), Nn);
    if (n_new_out) *n_new_out = Nn;
    return out;
}

// ── exact 3-NN mean distance (msplat_knn3_mean_dist, msplat_cpu.cpp:1499-1580) ─
// Brute force over a uniform grid: the three smallest squared distances (float, the
// reference's dx*dx + dy*dy + dz*dz order) are exact, so the mean of their square roots
// in ascending order equals the reference's k-d tree result.
void knn3(const float *pts, uint32_t n, float *out) {
    if (n == 0) return;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], pts[3 * i + a]);
            hi[a] = std::max(hi[a], pts[3 * i + a]);
        }
    const int G = std::max(1, (int)std::cbrt((double)n / 4.0));
    double cell[3];
    for (int a = 0; a < 3; a++) cell[a] = std::max(((double)hi[a] - lo[a]) / G, 1e-30);
    auto cid = [&](const float *p, int a) {
        int c = (int)(((double)p[a] - lo[a]) / cell[a]);
        return std::min(std::max(c, 0), G - 1);
    };
    std::vector<uint32_t> start((size_t)G * G * G + 1, 0), items(n);
    for (uint32_t i = 0; i < n; i++) start[(cid(pts + 3 * i, 2) * G + cid(pts + 3 * i, 1)) * G + cid(pts + 3 * i, 0) + 1]++;
    for (size_t c = 0; c < (size_t)G * G * G; c++) start[c + 1] += start[c];
    {
        std::vector<uint32_t> cur(start.begin(), start.end() - 1);
        for (uint32_t i = 0; i < n; i++) items[cur[(cid(pts + 3 * i, 2) * G + cid(pts + 3 * i, 1)) * G + cid(pts + 3 * i, 0)]++] = i;
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const uint32_t i = (uint32_t)ii;
        const float *q = pts + 3 * i;
        float best[3] = {INFINITY, INFINITY, INFINITY};
        auto insert = [&](float d) {
            if (d >= best[2]) return;
            if (d < best[1]) {
                best[2] = best[1];
                if (d < best[0]) { best[1] = best[0]; best[0] = d; }
                else best[1] = d;
            } else best[2] = d;
        };
        int c0[3] = {cid(q, 0), cid(q, 1), cid(q, 2)};
        for (int r = 0;; r++) {
            // visit the shell of cells at Chebyshev ring r
            for (int z = c0[2] - r; z <= c0[2] + r; z++)
                for (int y = c0[1] - r; y <= c0[1] + r; y++)
                    for (int x = c0[0] - r; x <= c0[0] + r; x++) {
                        if (std::max(std::abs(x - c0[0]), std::max(std::abs(y - c0[1]), std::abs(z - c0[2]))) != r) continue;
                        if (x < 0 || y < 0 || z < 0 || x >= G || y >= G || z >= G) continue;
                        size_t c = ((size_t)z * G + y) * G + x;
                        for (uint32_t s = start[c]; s < start[c + 1]; s++) {
                            uint32_t p = items[s];
                            if (p == i) continue;
                            float dx = pts[3 * p] - q[0], dy = pts[3 * p + 1] - q[1], dz = pts[3 * p + 2] - q[2];
                            insert(dx * dx + dy * dy + dz * dz);
                        }
                    }
            // every cell beyond ring r is at least r * min(cell) away
            double reach = r * std::min(cell[0], std::min(cell[1], cell[2]));
            if ((std::isfinite(best[2]) && reach * reach > (double)best[2] * (1 + 1e-6)) || r > G) break;
        }
        float sum = 0.f, cnt = 0.f;
        for (int k = 0; k < 3; k++)
            if (std::isfinite(best[k])) { sum += std::sqrt(best[k]); cnt += 1.f; }
        out[i] = cnt > 0.f ? sum / cnt : 1.f;
    }
}

// Gaussian init from points (msplat_init_gaussians, msplat_cpu.cpp:1582-1605)
void init_gaussians(const float *means, const uint8_t *rgb, uint32_t n, uint32_t seed, float logit, float *scales,
                    float *quats, float *dc, float *opac) {
    std::vector<float> md(n);
    knn3(means, n, md.data());
    const float tau = 6.28318530717958647692f;
    const float c0 = (float)kC0;
    for (uint32_t i = 0; i < n; i++) {
        float ls = std::log(std::max(md[i], 1e-7f));
        scales[3 * i] = scales[3 * i + 1] = scales[3 * i + 2] = ls;
        float u = uniform01(seed, 3u * i), v = uniform01(seed, 3u * i + 1u), w = uniform01(seed, 3u * i + 2u);
        quats[4 * i] = std::sqrt(1.f - u) * std::sin(tau * v);
        quats[4 * i + 1] = std::sqrt(1.f - u) * std::cos(tau * v);
        quats[4 * i + 2] = std::sqrt(u) * std::sin(tau * w);
        quats[4 * i + 3] = std::sqrt(u) * std::cos(tau * w);
        for (int k = 0; k < 3; k++) dc[3 * i + k] = ((float)rgb[3 * i + k] * (1.f / 255.f) - 0.5f) / c0;
        opac[i] = logit;
    }
}

// PSNR, clamp-to-edge SSIM, L1 (msplat_image_metrics, msplat_cpu.cpp:1671-1721)
void image_metrics(const float *a, const float *b, int h, int w, double *out) {
    out[0] = out[1] = out[2] = 0.0;
    if (h <= 0 || w <= 0) return;
    float G[11];
    for (int k = 0; k < 11; k++) G[k] = (float)kGauss[k];
    const float C1 = 0.0001f, C2 = 0.0009f;
    std::vector<float> mom((size_t)h * w * 15);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            for (int c = 0; c < 3; c++) {
                float s[5] = {0, 0, 0, 0, 0};
                for (int k = 0; k < 11; k++) {
                    int sx = std::min(std::max(x + k - 5, 0), w - 1);
                    size_t i = ((size_t)y * w + sx) * 3 + c;
                    float av = a[i], bv = b[i];
                    s[0] += G[k] * av; s[1] += G[k] * bv; s[2] += G[k] * av * av; s[3] += G[k] * bv * bv;
                    s[4] += G[k] * av * bv;
                }
                std::memcpy(&mom[((size_t)y * w + x) * 15 + c * 5], s, sizeof s);
            }
    double ssim = 0, l1 = 0, sq = 0;
    for (int y = 0; y < h; y++) {
        double rs = 0, rl = 0, rq = 0;
        for (int x = 0; x < w; x++)
            for (int c = 0; c < 3; c++) {
                float s[5] = {0, 0, 0, 0, 0};
                for (int k = 0; k < 11; k++) {
                    int sy = std::min(std::max(y + k - 5, 0), h - 1);
                    const float *o = &mom[((size_t)sy * w + x) * 15 + c * 5];
                    for (int q = 0; q < 5; q++) s[q] += G[k] * o[q];
                }
                float m12 = s[0] * s[1], m1 = s[0] * s[0], m2 = s[1] * s[1];
                float num = (2.f * m12 + C1) * (2.f * (s[4] - m12) + C2);
                float den = (m1 + m2 + C1) * ((s[2] - m1) + (s[3] - m2) + C2);
                rs += num / den;
                size_t i = ((size_t)y * w + x) * 3 + c;
                float d = a[i] - b[i];
                rl += std::fabs(d);
                rq += (double)d * d;
            }
        ssim += rs; l1 += rl; sq += rq;
    }
    double cnt = (double)h * w * 3, mse = sq / cnt;
    out[0] = mse > 0 ? 10.0 * std::log10(1.0 / mse) : INFINITY;
    out[1] = ssim / cnt;
    out[2] = l1 / cnt;
}

}  // namespace orc

// ═════════════════════════════ C entry points ══════════════════════════════
using namespace orc;

#define ORC_RENDER_API(T, S)                                                                                          \
    extern "C" void orc_##S##_project(int N, const T *means, const T *scales, T glob, const T *quats, const T *vm,    \
                                      const T *pm, T fx, T fy, T cx, T cy, uint32_t W, uint32_t H, int tx, int ty,     \
                                      T clip, uint32_t degree, uint32_t dtu, const T *cam, const T *dc,                \
                                      const T *rest, T *xys, T *depths, T *conics, T *colors, T *aabb,                  \
                                      int32_t *radii) {                                                                 \
        project<T>(N, means, scales, glob, quats, vm, pm, fx, fy, cx, cy, W, H, tx, ty, clip, degree, dtu, cam, dc,    \
                   rest, xys, depths, conics, colors, aabb, radii);                                                    \
    }                                                                                                                  \
    extern "C" int64_t orc_##S##_count_isects(int N, const T *xys, const T *aabb, const int32_t *radii, int tx,       \
                                              int ty, int32_t *tile_count) {                                           \
        return count_isects<T>(N, xys, aabb, radii, tx, ty, tile_count);                                               \
    }                                                                                                                  \
    extern "C" void orc_##S##_bin_sort(int N, const T *xys, const T *depths, const T *aabb, const int32_t *radii,     \
                                       const T *opac, const T *conics, const T *colors, int tx, int ty, int32_t *ts,   \
                                       int32_t *te, int32_t *gid, T *pk_xyo, T *pk_conic, T *pk_rgb,                   \
                                       uint64_t *keys) {                                                               \
        bin_sort<T>(N, xys, depths, aabb, radii, opac, conics, colors, tx, ty, ts, te, gid, pk_xyo, pk_conic, pk_rgb,  \
                    keys);                                                                                             \
    }                                                                                                                  \
    extern "C" void orc_##S##_raster_forward(uint32_t W, uint32_t H, int tx, int ty, const int32_t *ts,               \
                                             const int32_t *te, const T *pk_xyo, const T *pk_conic, const T *pk_rgb,   \
                                             const T *bg, T *out, T *final_T, int32_t *final_idx) {                    \
        raster_forward<T>(W, H, tx, ty, ts, te, pk_xyo, pk_conic, pk_rgb, bg, out, final_T, final_idx);                \
    }                                                                                                                  \
    extern "C" void orc_##S##_raster_backward(uint32_t W, uint32_t H, int tx, int ty, const int32_t *ts,              \
                                              const int32_t *te, const T *pk_xyo, const T *pk_conic,                   \
                                              const T *pk_rgb, const T *bg, const T *final_T,                          \
                                              const int32_t *final_idx, const T *v_out, T *slots) {                    \
        raster_backward<T>(W, H, tx, ty, ts, te, pk_xyo, pk_conic, pk_rgb, bg, final_T, final_idx, v_out, slots);      \
    }                                                                                                                  \
    extern "C" void orc_##S##_reduce_slots(int N, int64_t M, const int32_t *gid, const T *slots, T *v_xy,             \
                                           T *v_conic, T *v_rgb, T *v_opac, T *v_depth) {                              \
        reduce_slots<T>(N, M, gid, slots, v_xy, v_conic, v_rgb, v_opac, v_depth);                                      \
    }

ORC_RENDER_API(float, f32)
ORC_RENDER_API(double, f64)

extern "C" double orc_f32_loss(uint32_t W, uint32_t H, const float *ren, const float *gt, float w, float inv_n,
                               float *v_ren) {
    return loss_fwd_bwd(W, H, ren, gt, w, inv_n, v_ren);
}

extern "C" void orc_f32_project_backward_adam(
    int N, const float *means, const float *scales, float glob, const float *quats, const float *vm, const float *pm,
    float fx, float fy, uint32_t W, uint32_t H, uint32_t degree, uint32_t dtu, const float *cam, const int32_t *radii,
    const float *conics, const float *v_xy, const float *v_conic, const float *v_rgb, const float *v_depth,
    float *v_mean, float *v_scale, float *v_quat, float *dc, float *rest, float *dc_m1, float *dc_m2, float *rest_m1,
    float *rest_m2, float dc_step, float dc_bc2, float rest_step, float rest_bc2, float b1, float b2, float eps) {
    project_backward_adam<float>(N, means, scales, glob, quats, vm, pm, fx, fy, W, H, degree, dtu, cam, radii, conics,
                                 v_xy, v_conic, v_rgb, v_depth, v_mean, v_scale, v_quat, dc, rest, dc_m1, dc_m2,
                                 rest_m1, rest_m2, dc_step, dc_bc2, rest_step, rest_bc2, b1, b2, eps);
}

// The same projection backward in double (the FP64 renderer's parameter gradients: with
// b1 = b2 = 0 and zero step sizes the SH "Adam" leaves dc/rest unchanged and stores the
// raw gradients in dc_m1/rest_m1).
extern "C" void orc_f64_project_backward_adam(
    int N, const double *means, const double *scales, double glob, const double *quats, const double *vm,
    const double *pm, double fx, double fy, uint32_t W, uint32_t H, uint32_t degree, uint32_t dtu, const double *cam,
    const int32_t *radii, const double *conics, const double *v_xy, const double *v_conic, const double *v_rgb,
    const double *v_depth, double *v_mean, double *v_scale, double *v_quat, double *dc, double *rest, double *dc_m1,
    double *dc_m2, double *rest_m1, double *rest_m2, double dc_step, double dc_bc2, double rest_step, double rest_bc2,
    double b1, double b2, double eps) {
    project_backward_adam<double>(N, means, scales, glob, quats, vm, pm, fx, fy, W, H, degree, dtu, cam, radii,
                                  conics, v_xy, v_conic, v_rgb, v_depth, v_mean, v_scale, v_quat, dc, rest, dc_m1,
                                  dc_m2, rest_m1, rest_m2, dc_step, dc_bc2, rest_step, rest_bc2, b1, b2, eps);
}

extern "C" void orc_f32_adam(float *p, const float *g, float *m1, float *m2, size_t n, float step, float b1, float b2,
                             float bc2, float eps) {
    for (size_t i = 0; i < n; i++) adam<float>(p[i], m1[i], m2[i], g[i], step, b1, b2, bc2, eps);
}

extern "C" void orc_f32_grad_stats(int N, const int32_t *radii, const float *v_xy, float inv_max_dim, float *vis,
                                   float *gn, float *ms) {
    grad_stats<float>(N, radii, v_xy, inv_max_dim, vis, gn, ms);
}

// ── 9. prior-guided extras (msplat_cpu.cpp:327-402, 787-950, 1219-1309) ─────────
// PriorStep (core/metal/bindings.h:43-77) with plain host pointers; the HIP library's
// msplat_hip_prior has the same layout.
struct Prior {
    int32_t aux, sky;
    float *sky_tex, *sky_grad, *sky_m1, *sky_m2;
    int32_t sky_w, sky_h;
    float sky_frame[9];  // up, e1 (azimuth 0), e2 (azimuth 90 deg)
    float sky_step, sky_bc2;
    const float *prior_depth;
    const uint8_t *prior_aux;  // (prior_h, prior_w, 4): confidence, sky, keep, -
    int32_t prior_w, prior_h, has_depth, has_sky_mask;
    float depth_weight, sky_weight, fill_weight, huber_delta, min_alpha;
    int32_t mask_photometric;
    float *loss_terms;  // [4]: depth, sky, fill, -
    int32_t exposure, cam_index, num_cams;
    float *expo_params, *expo_m1, *expo_m2, *expo_grad;  // (num_cams, 12), (num_cams, 12) x 2, [12]
    float expo_step, expo_bc2, expo_reg;
    float log_max_scale_ratio;
};

// make_sky_params / pixel_world_dir / sky_uv / sky_taps / sky_sample (msplat_cpu.cpp:331-402)
struct Sky {
    float up[3] = {0, 0, 0}, e1[3] = {0, 0, 0}, e2[3] = {0, 0, 0};
    float fx = 1, fy = 1, cx = 0, cy = 0;
    uint32_t w = 1, h = 1;
    bool on = false;
};

inline Sky sky_of(const Prior *p, float fx, float fy, float cx, float cy) {
    Sky k;
    k.fx = fx; k.fy = fy; k.cx = cx; k.cy = cy;
    if (p && p->sky && p->sky_tex) {
        for (int i = 0; i < 3; i++) {
            k.up[i] = p->sky_frame[i];
            k.e1[i] = p->sky_frame[3 + i];
            k.e2[i] = p->sky_frame[6 + i];
        }
        k.w = (uint32_t)p->sky_w;
        k.h = (uint32_t)p->sky_h;
        k.on = true;
    }
    return k;
}

// world direction of the ray through pixel (px, py); vm is world-to-camera
inline v3<float> pixel_dir(float px, float py, const float *vm, const Sky &k) {
    const float dx = (px + 0.5f - k.cx) / k.fx, dy = (py + 0.5f - k.cy) / k.fy;
    return {vm[0] * dx + vm[4] * dy + vm[8], vm[1] * dx + vm[5] * dy + vm[9], vm[2] * dx + vm[6] * dy + vm[10]};
}

// equirect texel coordinates of direction d: azimuth around `up` from e1 towards e2
inline void sky_uv(v3<float> d, const Sky &k, float &u, float &v) {
    const float pi = 3.14159265358979323846f;
    d = mul(d, 1.f / std::sqrt(dot(d, d)));
    const v3<float> up{k.up[0], k.up[1], k.up[2]}, e1{k.e1[0], k.e1[1], k.e1[2]}, e2{k.e2[0], k.e2[1], k.e2[2]};
    const float el = std::asin(std::min(std::max(dot(d, up), -1.f), 1.f));
    const float az = std::atan2(dot(d, e2), dot(d, e1));
    u = (az * (0.5f / pi) + 0.5f) * (float)k.w - 0.5f;
    v = (0.5f - el / pi) * (float)k.h - 0.5f;
}

// bilinear taps: u wraps around, v clamps to the texture
inline void sky_taps(float u, float v, uint32_t w, uint32_t h, uint32_t idx[4], float wt[4]) {
    const float u0f = std::floor(u), fu = u - u0f;
    int u0 = (int)u0f % (int)w;
    if (u0 < 0) u0 += (int)w;
    const int u1 = (u0 + 1) % (int)w;
    const float vc = std::min(std::max(v, 0.f), (float)h - 1.f);
    const int v0 = (int)vc, v1 = std::min(v0 + 1, (int)h - 1);
    const float fv = vc - (float)v0;
    idx[0] = (uint32_t)(v0 * (int)w + u0); wt[0] = (1.f - fu) * (1.f - fv);
    idx[1] = (uint32_t)(v0 * (int)w + u1); wt[1] = fu * (1.f - fv);
    idx[2] = (uint32_t)(v1 * (int)w + u0); wt[2] = (1.f - fu) * fv;
    idx[3] = (uint32_t)(v1 * (int)w + u1); wt[3] = fu * fv;
}

// the sky colour behind every pixel (sky_sample of the pixel's ray), W x H x 3
void sky_image(uint32_t W, uint32_t H, const float *vm, const Sky &k, const float *tex, float *out) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < (int)H; y++)
        for (uint32_t x = 0; x < W; x++) {
            float u, v;
            sky_uv(pixel_dir((float)x, (float)y, vm, k), k, u, v);
            uint32_t idx[4];
            float wt[4];
            sky_taps(u, v, k.w, k.h, idx, wt);
            float c[3] = {0, 0, 0};
            for (int t = 0; t < 4; t++)
                for (int ch = 0; ch < 3; ch++) c[ch] = tex[3 * (size_t)idx[t] + ch] * wt[t] + c[ch];
            const size_t pid = (size_t)y * W + x;
            for (int ch = 0; ch < 3; ch++) out[3 * pid + ch] = c[ch];
        }
}

// sample_prior_depth (msplat_cpu.cpp:790-806): bilinear, but nearest across a depth edge
// (or next to an invalid sample), so that no surface is invented between two
inline float prior_depth_at(const float *d, float x, float y, uint32_t w, uint32_t h) {
    x = std::min(std::max(x, 0.f), (float)w - 1.f);
    y = std::min(std::max(y, 0.f), (float)h - 1.f);
    const uint32_t x0 = (uint32_t)x, y0 = (uint32_t)y;
    const uint32_t x1 = std::min(x0 + 1, w - 1), y1 = std::min(y0 + 1, h - 1);
    const float fx = x - (float)x0, fy = y - (float)y0;
    const float d00 = d[y0 * w + x0], d01 = d[y0 * w + x1], d10 = d[y1 * w + x0], d11 = d[y1 * w + x1];
    const float lo = std::min(std::min(d00, d01), std::min(d10, d11));
    const float hi = std::max(std::max(d00, d01), std::max(d10, d11));
    if (lo <= 0.f || hi > 1.05f * lo) return d[((fy < 0.5f) ? y0 : y1) * w + ((fx < 0.5f) ? x0 : x1)];
    const float top = d00 + (d01 - d00) * fx, bottom = d10 + (d11 - d10) * fx;
    return top + (bottom - top) * fy;
}

// prior_loss (msplat_cpu.cpp:809-884): per pixel dL/dD and dL/dA into v_aux, the
// photometric gradient zeroed where the prior says not to keep the pixel, and the depth,
// sky and fill terms (double per row, rows in order) added to loss_terms
void prior_loss(uint32_t W, uint32_t H, const Prior &p, const float *final_T, const float *out_depth, float *v_photo,
                float *v_aux) {
    const bool has_prior = p.prior_depth && p.prior_aux && p.prior_w > 0 && p.prior_h > 0;
    const uint32_t pw = has_prior ? (uint32_t)p.prior_w : 1u, ph = has_prior ? (uint32_t)p.prior_h : 1u;
    static const uint8_t kDefaultAux[4] = {0, 0, 255, 0};  // keep, not sky, no confidence
    static const float kDefaultDepth[1] = {0.f};
    const uint8_t *pa = has_prior ? p.prior_aux : kDefaultAux;
    const float *pd = has_prior ? p.prior_depth : kDefaultDepth;
    const bool has_depth = has_prior && p.has_depth, has_sky = has_prior && p.has_sky_mask;
    const float huber = std::max(p.huber_delta, 1e-6f);
    const float inv_npix = 1.0f / (float)(W * H);
    std::vector<double> terms((size_t)H * 3, 0.0);
#pragma omp parallel for schedule(static)
    for (int gy = 0; gy < (int)H; gy++) {
        double ld = 0, ls = 0, lf = 0;
        for (uint32_t gx = 0; gx < W; gx++) {
            const size_t pix = (size_t)gy * W + gx;
            const float sx = ((float)gx + 0.5f) * (float)pw / (float)W;
            const float sy = ((float)gy + 0.5f) * (float)ph / (float)H;
            const uint32_t q = std::min((uint32_t)sy, ph - 1) * pw + std::min((uint32_t)sx, pw - 1);
            const float conf = (float)pa[4 * q] * (1.f / 255.f);
            const bool is_sky = has_sky && pa[4 * q + 1] >= 128, keep = pa[4 * q + 2] >= 128;
            float v_d = 0.f, v_a = 0.f;
            if (!keep) {
                if (p.mask_photometric) v_photo[3 * pix] = v_photo[3 * pix + 1] = v_photo[3 * pix + 2] = 0.f;
            } else {
                const float A = 1.f - final_T[pix];
                if (is_sky) {
                    ls += p.sky_weight * A;
                    v_a = p.sky_weight * inv_npix;
                } else {
                    if (has_sky && p.fill_weight > 0.f) {
                        lf += p.fill_weight * (1.f - A);
                        v_a -= p.fill_weight * inv_npix;
                    }
                    const float D = out_depth[pix];
                    if (has_depth && p.depth_weight > 0.f && conf > 0.f && A > p.min_alpha && D > 0.f) {
                        const float dp = prior_depth_at(pd, sx - 0.5f, sy - 0.5f, pw, ph);
                        if (dp > 0.f) {
                            const float r = std::log(D / A) - std::log(dp), ar = std::fabs(r);
                            float dr;
                            if (ar < huber) {
                                ld += p.depth_weight * conf * 0.5f * r * r / huber;
                                dr = r / huber;
                            } else {
                                ld += p.depth_weight * conf * (ar - 0.5f * huber);
                                dr = r > 0.f ? 1.f : (r < 0.f ? -1.f : 0.f);
                            }
                            const float sc = p.depth_weight * conf * dr * inv_npix;
                            v_d = sc / D;
                            v_a -= sc / A;
                        }
                    }
                }
            }
            v_aux[2 * pix] = v_d;
            v_aux[2 * pix + 1] = v_a;
        }
        terms[3 * (size_t)gy] = ld;
        terms[3 * (size_t)gy + 1] = ls;
        terms[3 * (size_t)gy + 2] = lf;
    }
    if (p.loss_terms) {
        double t[3] = {0, 0, 0};
        for (uint32_t y = 0; y < H; y++)
            for (int c = 0; c < 3; c++) t[c] += terms[3 * (size_t)y + c];
        for (int c = 0; c < 3; c++) p.loss_terms[c] += (float)(t[c] * inv_npix);
    }
}

// exposure_apply (msplat_cpu.cpp:887-896): rows 0-2 of P are the 3x3 colour matrix, 9-11 the offset
void exposure_apply(const float *img, const float *P, size_t n, float *out) {
    for (size_t i = 0; i < n; i++) {
        const float c0 = img[3 * i], c1 = img[3 * i + 1], c2 = img[3 * i + 2];
        out[3 * i] = P[0] * c0 + P[1] * c1 + P[2] * c2 + P[9];
        out[3 * i + 1] = P[3] * c0 + P[4] * c1 + P[5] * c2 + P[10];
        out[3 * i + 2] = P[6] * c0 + P[7] * c1 + P[8] * c2 + P[11];
    }
}

// exposure_backward_adam (msplat_cpu.cpp:899-924): the render's gradient P^T v_adj, the
// 12 parameter gradients (double sums over the pixels in order) plus the caller's extra
// gradient and the pull towards the identity, one Adam step on camera cam_index's row
void exposure_backward_adam(const float *img, const Prior &p, size_t n, const float *v_adj, float *v_img, float b1,
                            float b2, float eps) {
    const uint32_t cam = (uint32_t)std::max(0, p.cam_index);
    float *params = p.expo_params;
    const float *P = params + 12 * cam;
    double g[12] = {0};
    for (size_t i = 0; i < n; i++) {
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
    for (int k = 0; k < 12; k++) {
        const uint32_t i = 12 * cam + k;
        const float ident = (k == 0 || k == 4 || k == 8) ? 1.f : 0.f;
        const float gk = (float)g[k] + (p.expo_grad ? p.expo_grad[k] : 0.f) + p.expo_reg * (params[i] - ident);
        if (p.expo_grad) p.expo_grad[k] = 0.f;
        adam<float>(params[i], p.expo_m1[i], p.expo_m2[i], gk, p.expo_step, b1, b2, p.expo_bc2, eps);
    }
}

// sky_backward_adam (msplat_cpu.cpp:927-950): each pixel with T >= T_MIN adds T v_img to
// its ray's four taps (pixels in row order), then one Adam step on every texel
void sky_backward_adam(uint32_t W, uint32_t H, const Prior &p, const Sky &k, const float *vm, const float *final_T,
                       const float *v_img, float b1, float b2, float eps) {
    float *grad = p.sky_grad;
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            const size_t pix = (size_t)y * W + x;
            const float T = final_T[pix];
            if (T < 1e-4f) continue;
            const float g[3] = {T * v_img[3 * pix], T * v_img[3 * pix + 1], T * v_img[3 * pix + 2]};
            float u, v;
            sky_uv(pixel_dir((float)x, (float)y, vm, k), k, u, v);
            uint32_t idx[4];
            float wt[4];
            sky_taps(u, v, k.w, k.h, idx, wt);
            for (int t = 0; t < 4; t++) {
                if (wt[t] == 0.f) continue;
                float *d = grad + 3 * (size_t)idx[t];
                d[0] += wt[t] * g[0];
                d[1] += wt[t] * g[1];
                d[2] += wt[t] * g[2];
            }
        }
    const size_t n = (size_t)p.sky_w * p.sky_h * 3;
    for (size_t i = 0; i < n; i++) adam<float>(p.sky_tex[i], p.sky_m1[i], p.sky_m2[i], grad[i], p.sky_step, b1, b2, p.sky_bc2, eps);
}

// scale_ratio_cap (msplat_cpu.cpp:1295-1309): a gaussian's largest log-scale is lowered to
// its middle one plus log_max (the largest of equal values: the first)
void scale_cap(int N, float *s, float log_max) {
    for (int i = 0; i < N; i++) {
        const float a = s[3 * i], b = s[3 * i + 1], c = s[3 * i + 2];
        const float mx = std::max(a, std::max(b, c)), mn = std::min(a, std::min(b, c));
        const float cap = (a + b + c - mx - mn) + log_max;
        if (mx <= cap) continue;
        if (a == mx) s[3 * i] = cap;
        else if (b == mx) s[3 * i + 1] = cap;
        else s[3 * i + 2] = cap;
    }
}

// One training step (msplat_train_step, msplat_cpu.cpp:1204-1327), with the
// prior-guided extras when prior is given (aux forward with the learned sky, exposure on
// the photometric loss, prior losses, sky and exposure Adam, aux backward, needle cap).
// params/m1/m2 hold the six Adam groups (means, scales, quats, dc, rest, opacity).
// Returns the mean per-pixel photometric loss; radii_out receives the radii, img_out the
// render, depth_out / final_T_out (aux) the expected-depth sum and final transmittance.
static float train_step_f32(int N, float glob, const float *vm, const float *pm, float fx, float fy, float cx,
                            float cy, uint32_t H, uint32_t W, int tx, int ty, float clip, uint32_t degree,
                            uint32_t dtu, const float *cam, const float *bg, const float *gt, float ssim_weight,
                            float loss_inv_n, float **params, float **m1, float **m2, const float *steps,
                            const float *bc2s, float b1, float b2, float eps, float *vis, float *gn, float *ms,
                            float inv_max_dim, int32_t *radii_out, float *img_out, const Prior *prior,
                            float *depth_out, float *final_T_out) {
    const bool aux = prior && prior->aux;
    const bool sky = aux && prior->sky && prior->sky_tex && prior->sky_grad;
    const bool exposure = prior && prior->exposure && prior->expo_params;
    const bool cap = prior && prior->log_max_scale_ratio > 0.f;
    std::vector<float> xys(2 * (size_t)N), depths(N), conics(3 * (size_t)N), colors(3 * (size_t)N), aabb(2 * (size_t)N);
    std::vector<int32_t> radii(N);
    project<float>(N, params[0], params[1], glob, params[2], vm, pm, fx, fy, cx, cy, W, H, tx, ty, clip, degree, dtu,
                   cam, params[3], params[4], xys.data(), depths.data(), conics.data(), colors.data(), aabb.data(),
                   radii.data());
    const int nt = tx * ty;
    int64_t M = count_isects<float>(N, xys.data(), aabb.data(), radii.data(), tx, ty, nullptr);
    const size_t Mc = (size_t)std::max<int64_t>(M, 1);
    std::vector<int32_t> ts(nt), te(nt), gid(Mc);
    std::vector<float> pxyo(3 * Mc), pcn(3 * Mc), prgb(3 * Mc);
    bin_sort<float>(N, xys.data(), depths.data(), aabb.data(), radii.data(), params[5], conics.data(), colors.data(),
                    tx, ty, ts.data(), te.data(), gid.data(), pxyo.data(), pcn.data(), prgb.data(), nullptr);
    const size_t P = (size_t)W * H;
    std::vector<float> img(3 * P), fT(P), vimg(3 * P);
    std::vector<int32_t> fidx(P);
    std::vector<float> depth, bg_img, sky_bg, v_aux;
    AuxFwd<float> af;
    const Sky sky_fwd = sky_of(aux ? prior : nullptr, fx, fy, cx, cy);
    if (aux) {
        depth.assign(P, 0.f);
        bg_img.assign(3 * P, 0.f);
        af.depths = depths.data();
        af.gid = gid.data();
        af.out_depth = depth.data();
        af.bg_img = bg_img.data();
        if (sky_fwd.on) {
            sky_bg.resize(3 * P);
            sky_image(W, H, vm, sky_fwd, prior->sky_tex, sky_bg.data());
            af.sky_bg = sky_bg.data();
        }
    }
    raster_forward<float>(W, H, tx, ty, ts.data(), te.data(), pxyo.data(), pcn.data(), prgb.data(), bg, img.data(),
                          fT.data(), fidx.data(), aux ? &af : nullptr);
    if (aux && prior->loss_terms) std::fill(prior->loss_terms, prior->loss_terms + 4, 0.f);
    if (sky) std::fill(prior->sky_grad, prior->sky_grad + (size_t)prior->sky_w * prior->sky_h * 3, 0.f);
    // photometric loss, on the exposure-adjusted render with exposure
    const float *loss_in = img.data();
    float *loss_grad = vimg.data();
    std::vector<float> adj, v_adj;
    if (exposure) {
        adj.resize(3 * P);
        v_adj.resize(3 * P);
        exposure_apply(img.data(), prior->expo_params + 12 * std::max(0, prior->cam_index), P, adj.data());
        loss_in = adj.data();
        loss_grad = v_adj.data();
    }
    double lsum = loss_fwd_bwd(W, H, loss_in, gt, ssim_weight, loss_inv_n, loss_grad);
    if (aux) {
        v_aux.resize(2 * P);
        prior_loss(W, H, *prior, fT.data(), depth.data(), loss_grad, v_aux.data());
    }
    if (exposure) exposure_backward_adam(img.data(), *prior, P, v_adj.data(), vimg.data(), b1, b2, eps);
    if (sky)
        sky_backward_adam(W, H, *prior, sky_of(prior, fx, fy, cx, cy), vm, fT.data(), vimg.data(), b1, b2, eps);
    std::vector<float> slots(10 * Mc);
    AuxBwd<float> ab;
    if (aux) {
        ab.depths = depths.data();
        ab.gid = gid.data();
        ab.bg_img = bg_img.data();
        ab.v_aux = v_aux.data();
    }
    raster_backward<float>(W, H, tx, ty, ts.data(), te.data(), pxyo.data(), pcn.data(), prgb.data(), bg, fT.data(),
                           fidx.data(), vimg.data(), slots.data(), aux ? &ab : nullptr);
    std::vector<float> vxy(2 * (size_t)N), vcn(3 * (size_t)N), vrgb(3 * (size_t)N), vop(N), vdp(N);
    reduce_slots<float>(N, M, gid.data(), slots.data(), vxy.data(), vcn.data(), vrgb.data(), vop.data(), vdp.data());
    std::vector<float> vmean(3 * (size_t)N), vsc(3 * (size_t)N), vq(4 * (size_t)N);
    project_backward_adam<float>(N, params[0], params[1], glob, params[2], vm, pm, fx, fy, W, H, degree, dtu, cam,
                                 radii.data(), conics.data(), vxy.data(), vcn.data(), vrgb.data(), vdp.data(),
                                 vmean.data(), vsc.data(), vq.data(), params[3], params[4], m1[3], m2[3], m1[4], m2[4],
                                 steps[3], bc2s[3], steps[4], bc2s[4], b1, b2, eps);
    const float *grads[6] = {vmean.data(), vsc.data(), vq.data(), nullptr, nullptr, vop.data()};
    const size_t width[6] = {3, 3, 4, 0, 0, 1};
    for (int g : {0, 1, 2, 5})
        for (size_t i = 0; i < (size_t)N * width[g]; i++)
            adam<float>(params[g][i], m1[g][i], m2[g][i], grads[g][i], steps[g], b1, b2, bc2s[g], eps);
    if (cap) scale_cap(N, params[1], prior->log_max_scale_ratio);
    grad_stats<float>(N, radii.data(), vxy.data(), inv_max_dim, vis, gn, ms);
    if (radii_out) std::memcpy(radii_out, radii.data(), sizeof(int32_t) * N);
    if (img_out) std::memcpy(img_out, img.data(), sizeof(float) * 3 * P);
    if (depth_out && aux) std::memcpy(depth_out, depth.data(), sizeof(float) * P);
    if (final_T_out) std::memcpy(final_T_out, fT.data(), sizeof(float) * P);
    return (float)(lsum / (double)P);
}

extern "C" float orc_f32_train_step(int N, float glob, const float *vm, const float *pm, float fx, float fy, float cx,
                                    float cy, uint32_t H, uint32_t W, int tx, int ty, float clip, uint32_t degree,
                                    uint32_t dtu, const float *cam, const float *bg, const float *gt,
                                    float ssim_weight, float loss_inv_n, float **params, float **m1, float **m2,
                                    const float *steps, const float *bc2s, float b1, float b2, float eps, float *vis,
                                    float *gn, float *ms, float inv_max_dim, int32_t *radii_out, float *img_out) {
    return train_step_f32(N, glob, vm, pm, fx, fy, cx, cy, H, W, tx, ty, clip, degree, dtu, cam, bg, gt, ssim_weight,
                          loss_inv_n, params, m1, m2, steps, bc2s, b1, b2, eps, vis, gn, ms, inv_max_dim, radii_out,
                          img_out, nullptr, nullptr, nullptr);
}

// msplat_train_step with a PriorStep (msplat_cpu.cpp:1204-1327)
extern "C" float orc_f32_train_step_prior(int N, float glob, const float *vm, const float *pm, float fx, float fy,
                                          float cx, float cy, uint32_t H, uint32_t W, int tx, int ty, float clip,
                                          uint32_t degree, uint32_t dtu, const float *cam, const float *bg,
                                          const float *gt, float ssim_weight, float loss_inv_n, float **params,
                                          float **m1, float **m2, const float *steps, const float *bc2s, float b1,
                                          float b2, float eps, float *vis, float *gn, float *ms, float inv_max_dim,
                                          int32_t *radii_out, float *img_out, const Prior *prior, float *depth_out,
                                          float *final_T_out) {
    return train_step_f32(N, glob, vm, pm, fx, fy, cx, cy, H, W, tx, ty, clip, degree, dtu, cam, bg, gt, ssim_weight,
                          loss_inv_n, params, m1, m2, steps, bc2s, b1, b2, eps, vis, gn, ms, inv_max_dim, radii_out,
                          img_out, prior, depth_out, final_T_out);
}

// msplat_render with a PriorStep (msplat_cpu.cpp:1183-1197, 631-657): the aux forward
// with the learned sky; depth_out / final_T_out as msplat_render_aux_outputs
extern "C" void orc_f32_render_prior(int N, float glob, const float *vm, const float *pm, float fx, float fy, float cx,
                                     float cy, uint32_t H, uint32_t W, int tx, int ty, float clip, uint32_t degree,
                                     uint32_t dtu, const float *cam, float **params, const float *bg,
                                     const Prior *prior, float *img_out, float *depth_out, float *final_T_out,
                                     int32_t *radii_out) {
    const bool aux = prior && prior->aux;
    std::vector<float> xys(2 * (size_t)N), depths(N), conics(3 * (size_t)N), colors(3 * (size_t)N), aabb(2 * (size_t)N);
    std::vector<int32_t> radii(N);
    project<float>(N, params[0], params[1], glob, params[2], vm, pm, fx, fy, cx, cy, W, H, tx, ty, clip, degree, dtu,
                   cam, params[3], params[4], xys.data(), depths.data(), conics.data(), colors.data(), aabb.data(),
                   radii.data());
    const int nt = tx * ty;
    int64_t M = count_isects<float>(N, xys.data(), aabb.data(), radii.data(), tx, ty, nullptr);
    const size_t Mc = (size_t)std::max<int64_t>(M, 1);
    std::vector<int32_t> ts(nt), te(nt), gid(Mc), fidx((size_t)W * H);
    std::vector<float> pxyo(3 * Mc), pcn(3 * Mc), prgb(3 * Mc), fT((size_t)W * H);
    bin_sort<float>(N, xys.data(), depths.data(), aabb.data(), radii.data(), params[5], conics.data(), colors.data(),
                    tx, ty, ts.data(), te.data(), gid.data(), pxyo.data(), pcn.data(), prgb.data(), nullptr);
    const size_t P = (size_t)W * H;
    std::vector<float> depth(P, 0.f), bg_img(3 * P), sky_bg;
    AuxFwd<float> af;
    af.depths = depths.data();
    af.gid = gid.data();
    af.out_depth = depth.data();
    af.bg_img = bg_img.data();
    const Sky k = sky_of(aux ? prior : nullptr, fx, fy, cx, cy);
    if (k.on) {
        sky_bg.resize(3 * P);
        sky_image(W, H, vm, k, prior->sky_tex, sky_bg.data());
        af.sky_bg = sky_bg.data();
    }
    raster_forward<float>(W, H, tx, ty, ts.data(), te.data(), pxyo.data(), pcn.data(), prgb.data(), bg, img_out,
                          fT.data(), fidx.data(), aux ? &af : nullptr);
    if (depth_out && aux) std::memcpy(depth_out, depth.data(), sizeof(float) * P);
    if (final_T_out) std::memcpy(final_T_out, fT.data(), sizeof(float) * P);
    if (radii_out) std::memcpy(radii_out, radii.data(), sizeof(int32_t) * N);
}

// msplat_finalize_depth (msplat_cpu.cpp:1736-1745)
extern "C" void orc_finalize_depth(const float *d, const float *t, uint32_t n, float inv_scale, float *od, float *oa) {
    for (uint32_t i = 0; i < n; i++) {
        const float a = 1.f - t[i];
        oa[i] = a;
        od[i] = a > 1e-4f ? d[i] / a * inv_scale : 0.f;
    }
}

extern "C" int orc_densify(int N, int cap, float grad_thr, float size_thr, float screen_thr, int check_screen,
                           float cull_alpha, float cull_scale, float cull_screen, int check_huge, const float *gn,
                           const float *vc, const float *ms, float half_max_dim, float **bufs, int fr_stride,
                           uint32_t seed, uint8_t *split_out, uint8_t *dup_out, uint8_t *keep_out, int *n_new) {
    return densify(N, cap, grad_thr, size_thr, screen_thr, check_screen, cull_alpha, cull_scale, cull_screen,
                   check_huge, gn, vc, ms, half_max_dim, bufs, fr_stride, seed, split_out, dup_out, keep_out, n_new);
}

// msplat_opacity_reset, msplat_cpu.cpp:1467-1475
extern "C" void orc_opacity_reset(float *o, float *m1, float *m2, int n, float reset_logit) {
    for (int i = 0; i < n; i++) {
        o[i] = std::min(o[i], reset_logit);
        m1[i] = 0.f;
        m2[i] = 0.f;
    }
}

// Stable sort by the low key_bits bits (msplat_radix_sort, msplat_cpu.cpp:1481-1494)
extern "C" void orc_radix_sort(uint64_t *keys, int32_t *vals, uint32_t n, int key_bits) {
    if (n == 0) return;
    const uint64_t mask = key_bits >= 64 ? ~0ull : ((1ull << key_bits) - 1ull);
    std::vector<uint32_t> ord(n);
    std::iota(ord.begin(), ord.end(), 0u);
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return (keys[a] & mask) < (keys[b] & mask); });
    std::vector<uint64_t> k2(n);
    std::vector<int32_t> v2(n);
    for (uint32_t i = 0; i < n; i++) { k2[i] = keys[ord[i]]; v2[i] = vals[ord[i]]; }
    std::memcpy(keys, k2.data(), n * 8);
    std::memcpy(vals, v2.data(), n * 4);
}

extern "C" void orc_knn3_mean_dist(const float *pts, uint32_t n, float *out) { knn3(pts, n, out); }

extern "C" void orc_init_gaussians(const float *means, const uint8_t *rgb, uint32_t n, uint32_t seed, float logit,
                                   float *scales, float *quats, float *dc, float *opac) {
    init_gaussians(means, rgb, n, seed, logit, scales, quats, dc, opac);
}

extern "C" void orc_image_metrics(const float *a, const float *b, int h, int w, double *out) {
    image_metrics(a, b, h, w, out);
}

// RGBA8 packing by truncation (msplat_pack_rgba8, msplat_cpu.cpp:1723-1734)
extern "C" void orc_pack_rgba8(const float *img, uint32_t n, uint8_t *out) {
    for (uint32_t i = 0; i < n; i++) {
        for (int c = 0; c < 3; c++) out[4 * i + c] = (uint8_t)(std::min(std::max(img[3 * i + c], 0.f), 1.f) * 255.f);
        out[4 * i + 3] = 255;
    }
}

// msplat_resize_area (msplat_cpu.cpp:1609-1639): box resampling; every output pixel
// averages the source pixels under its footprint, weighted by overlap area (rows outer,
// columns inner). u8 sources are RGBA8 scaled by 1/255.
extern "C" void orc_resize_area(const void *src, int src_is_u8, int sw, int sh, float *out, int dw, int dh) {
    if (dw <= 0 || dh <= 0) return;
    const uint8_t *u8 = src_is_u8 ? (const uint8_t *)src : nullptr;
    const float *f = src_is_u8 ? nullptr : (const float *)src;
    const float scx = (float)sw / (float)dw, scy = (float)sh / (float)dh;
#pragma omp parallel for schedule(static)
    for (int gy = 0; gy < dh; gy++) {
        const float y0 = (float)gy * scy, y1 = (float)(gy + 1) * scy;
        const int iy0 = (int)y0, iy1 = std::min((int)std::ceil(y1), sh);
        for (int gx = 0; gx < dw; gx++) {
            const float x0 = (float)gx * scx, x1 = (float)(gx + 1) * scx;
            const int ix0 = (int)x0, ix1 = std::min((int)std::ceil(x1), sw);
            float acc[3] = {0.f, 0.f, 0.f}, area = 0.f;
            for (int iy = iy0; iy < iy1; iy++) {
                const float wy = std::min((float)(iy + 1), y1) - std::max((float)iy, y0);
                for (int ix = ix0; ix < ix1; ix++) {
                    const float wx = std::min((float)(ix + 1), x1) - std::max((float)ix, x0);
                    const size_t si = (size_t)iy * sw + ix;
                    const float w = wx * wy;
                    for (int c = 0; c < 3; c++) {
                        const float v = u8 ? (float)u8[4 * si + c] * (1.f / 255.f) : f[3 * si + c];
                        acc[c] = acc[c] + w * v;
                    }
                    area = area + w;
                }
            }
            for (int c = 0; c < 3; c++) out[3 * ((size_t)gy * dw + gx) + c] = acc[c] / area;
        }
    }
}

// msplat_undistort (msplat_cpu.cpp:1641-1669): Brown-Conrady model from the undistorted
// crop's pixel centres to the source, bilinear sampling with clamp-to-edge.
extern "C" void orc_undistort(const float *in, int sw, int sh, const float *intr, const float *dist, int roi_x,
                              int roi_y, float *out, int dw, int dh) {
    if (dw <= 0 || dh <= 0) return;
    const float fx = intr[0], fy = intr[1], cx = intr[2], cy = intr[3];
    const float k1 = dist[0], k2 = dist[1], p1 = dist[2], p2 = dist[3], k3 = dist[4];
#pragma omp parallel for schedule(static)
    for (int gy = 0; gy < dh; gy++)
        for (int gx = 0; gx < dw; gx++) {
            const float x = ((float)(gx + roi_x) + 0.5f - cx) / fx, y = ((float)(gy + roi_y) + 0.5f - cy) / fy;
            const float r2 = x * x + y * y;
            const float radial = 1.f + k1 * r2 + k2 * r2 * r2 + k3 * r2 * r2 * r2;
            const float xd = x * radial + 2.f * p1 * x * y + p2 * (r2 + 2.f * x * x);
            const float yd = y * radial + p1 * (r2 + 2.f * y * y) + 2.f * p2 * x * y;
            const float sx = xd * fx + cx - 0.5f, sy = yd * fy + cy - 0.5f;
            const int x0 = (int)std::floor(sx), y0 = (int)std::floor(sy);
            const float ax = sx - std::floor(sx), ay = sy - std::floor(sy);
            const int xa = std::clamp(x0, 0, sw - 1), xb = std::clamp(x0 + 1, 0, sw - 1);
            const int ya = std::clamp(y0, 0, sh - 1), yb = std::clamp(y0 + 1, 0, sh - 1);
            for (int c = 0; c < 3; c++) {
                const float v00 = in[3 * ((size_t)ya * sw + xa) + c], v10 = in[3 * ((size_t)ya * sw + xb) + c];
                const float v01 = in[3 * ((size_t)yb * sw + xa) + c], v11 = in[3 * ((size_t)yb * sw + xb) + c];
                const float top = v00 * (1.f - ax) + v10 * ax, bottom = v01 * (1.f - ax) + v11 * ax;
                out[3 * ((size_t)gy * dw + gx) + c] = top * (1.f - ay) + bottom * ay;
            }
        }
}

extern "C" uint32_t orc_pcg_hash(uint32_t v) { return pcg(v); }
extern "C" float orc_rand_uniform(uint32_t seed, uint32_t ctr) { return uniform01(seed, ctr); }
extern "C" void orc_rand_normal3(uint32_t seed, uint32_t stream, float *out) { normal3(seed, stream, out); }

// ── Host-side schedule math of the reference's Model (model.cpp), restated in float ──

// The integer part of Model::prepareCam (model.cpp:498-546) plus levelSize
// (image_sizing.hpp:62): pyramid size for downscale factor sf, intrinsics scaled by the
// per-axis ratio, fov from glibc atanf, viewmat = inverse of the OpenGL->OpenCV-flipped
// camToWorld, projmat = P(znear 1e-3, zfar 1e3) * viewmat in float, left-to-right sums.
extern "C" void orc_prepare_cam(const float *c2w, int W, int H, float fx, float fy, float cx, float cy, int sf,
                                float *viewmat, float *projmat, float *cam_pos, float *intr, int *wh) {
    const int w = sf <= 1 ? W : W / sf, h = sf <= 1 ? H : H / sf;
    const float rx = W > 0 ? (float)w / (float)W : 1.f / (float)sf;
    const float ry = H > 0 ? (float)h / (float)H : 1.f / (float)sf;
    const float sfx = fx * rx, scx = cx * rx, sfy = fy * ry, scy = cy * ry;
    const float fovX = 2.0f * std::atan(w / (2.0f * sfx));
    const float fovY = 2.0f * std::atan(h / (2.0f * sfy));
    float R[3][3], Ri[3][3], T[3], Ti[3];
    for (int i = 0; i < 3; i++) {
        R[i][0] = c2w[i * 4 + 0];
        R[i][1] = -c2w[i * 4 + 1];
        R[i][2] = -c2w[i * 4 + 2];
        T[i] = c2w[i * 4 + 3];
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Ri[i][j] = R[j][i];
    for (int i = 0; i < 3; i++) Ti[i] = -(Ri[i][0] * T[0] + Ri[i][1] * T[1] + Ri[i][2] * T[2]);
    const float vm[16] = {Ri[0][0], Ri[0][1], Ri[0][2], Ti[0], Ri[1][0], Ri[1][1], Ri[1][2], Ti[1],
                          Ri[2][0], Ri[2][1], Ri[2][2], Ti[2], 0.f,      0.f,      0.f,      1.f};
    const float zn = 0.001f, zf = 1000.0f;
    const float t_p = zn * std::tan(0.5f * fovY), r_p = zn * std::tan(0.5f * fovX);
    const float pm[16] = {zn / r_p, 0, 0, 0, 0, zn / t_p, 0, 0, 0, 0, (zf + zn) / (zf - zn), -zf * zn / (zf - zn),
                          0, 0, 1, 0};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float acc = 0.f;
            for (int k = 0; k < 4; k++) acc += pm[i * 4 + k] * vm[k * 4 + j];
            projmat[i * 4 + j] = acc;
        }
    std::memcpy(viewmat, vm, sizeof vm);
    for (int i = 0; i < 3; i++) cam_pos[i] = T[i];
    intr[0] = sfx;
    intr[1] = sfy;
    intr[2] = scx;
    intr[3] = scy;
    wh[0] = w;
    wh[1] = h;
}

// Model::schedulersStep (model.cpp:140-143): the means learning rate after `step`,
// exp(lerp(log init, log final, clamp(step / max_steps))) in float (glibc expf/logf).
extern "C" float orc_means_lr(int step, int max_steps, float lr_init, float lr_final) {
    const float t = std::clamp((float)step / (float)max_steps, 0.f, 1.f);
    return std::exp(std::log(lr_init) * (1.f - t) + std::log(lr_final) * t);
}

// Model::fullIteration's Adam scalars (model.cpp:638-651): bc = 1 - pow(beta, t) with
// the float beta promoted to double by std::pow(float, int), rounded to float; the step
// size lr / bc1 and sqrt(bc2) in float.
extern "C" void orc_adam_scalars(float lr, float beta1, float beta2, int t, float *step_size, float *bc2_sqrt) {
    const float bc1 = 1.0f - std::pow(beta1, t);
    const float bc2 = 1.0f - std::pow(beta2, t);
    *step_size = lr / bc1;
    *bc2_sqrt = std::sqrt(bc2);
}

// Model::afterTrain's cadence (model.cpp:199-245): bit 0 = densify at this step, bit 1
// = opacity reset, bit 2 = statistics restart, bit 3 = screen-size check on, bit 4 =
// huge-cull on.
extern "C" int orc_refine_flags(int step, int refine_every, int warmup, int reset_alpha_every, int stop_split_at,
                                int num_cameras, int stop_screen_size_at) {
    if (step % refine_every != 0 || step <= warmup) return 0;
    const int reset_interval = reset_alpha_every * refine_every;
    int f = 4;
    if (step < stop_split_at && step % reset_interval > num_cameras + refine_every) {
        f |= 1;
        if (step < stop_screen_size_at) f |= 8;
        if (step > refine_every * reset_alpha_every) f |= 16;
    }
    if (step < stop_split_at && step % reset_interval == refine_every) f |= 2;
    return f;
}
