// Plain-C entry points over the reference CPU backend (core/metal/bindings.h as
// implemented by core/cpu/msplat_cpu.cpp), so that Python can call the compiled
// reference with numpy buffers. TEST INFRASTRUCTURE ONLY: built by oracle/Makefile
// into oracle/_ref/ where /root/reference exists, used to write the golden vectors and
// as the CPU baseline; never part of the product.
#include "bindings.h"
#include "loaders.hpp"

#include <cstdio>
#include <cstring>
#include <exception>
#include <vector>

namespace {

MTensor upload(const void *src, std::vector<int64_t> shape, DType dt) {
    MTensor t = gpu_empty(std::move(shape), dt);
    if (src) std::memcpy(t.data_ptr(), src, t.nbytes());
    return t;
}

void download(void *dst, const MTensor &t) {
    if (dst) std::memcpy(dst, t.data_ptr(), t.nbytes());
}

int fail(const std::exception &e) {
    std::fprintf(stderr, "msplat reference: %s\n", e.what());
    return -1;
}

struct Scene {
    MTensor means, scales, quats, dc, rest, opac, vm, pm, bg;
    Scene(int N, const float *m, const float *s, const float *q, const float *d, const float *r, int rest_bases,
          const float *o, const float *viewmat, const float *projmat, const float *background) {
        means = upload(m, {N, 3}, DType::Float32);
        scales = upload(s, {N, 3}, DType::Float32);
        quats = upload(q, {N, 4}, DType::Float32);
        dc = upload(d, {N, 3}, DType::Float32);
        rest = upload(r, {N, rest_bases > 0 ? rest_bases : 1, 3}, DType::Float32);
        opac = upload(o, {N, 1}, DType::Float32);
        vm = upload(viewmat, {4, 4}, DType::Float32);
        pm = upload(projmat, {4, 4}, DType::Float32);
        bg = upload(background, {3}, DType::Float32);
    }
};

std::tuple<int, int, int> tiles(int W, int H) { return std::make_tuple((W + 15) / 16, (H + 15) / 16, 1); }

}  // namespace

extern "C" int ref_render(int N, const float *means, const float *scales, const float *quats, const float *dc,
                          const float *rest, int rest_bases, const float *opac, const float *viewmat,
                          const float *projmat, float fx, float fy, float cx, float cy, int H, int W, float clip,
                          int degree, int dtu, const float *cam_pos, const float *bg, float *out_img,
                          float *final_T) {
    try {
        Scene s(N, means, scales, quats, dc, rest, rest_bases, opac, viewmat, projmat, bg);
        float cam[3] = {cam_pos[0], cam_pos[1], cam_pos[2]};
        MTensor img = msplat_render(N, s.means, s.scales, 1.0f, s.quats, s.vm, s.pm, fx, fy, cx, cy, (unsigned)H,
                                    (unsigned)W, tiles(W, H), clip, (unsigned)degree, (unsigned)dtu, cam, s.dc, s.rest,
                                    s.opac, s.bg, nullptr);
        msplat_gpu_sync();
        download(out_img, img);
        if (final_T) {
            MTensor depth, fT;
            msplat_render_aux_outputs(depth, fT);
            download(final_T, fT);
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

// One msplat_train_step. params[6] and m1[6]/m2[6] are read and written in place
// (means, scales, quats, dc, rest, opacity); vis/gn/ms are accumulated in place.
extern "C" int ref_train_step(int N, float **params, int rest_bases, float **m1, float **m2, const float *viewmat,
                              const float *projmat, float fx, float fy, float cx, float cy, int H, int W, float clip,
                              int degree, int dtu, const float *cam_pos, const float *bg, const float *gt,
                              float ssim_weight, const float *steps, const float *bc2s, float b1, float b2, float eps,
                              float *vis, float *gn, float *ms, int32_t *radii_out, float *loss_out) {
    try {
        Scene s(N, params[0], params[1], params[2], params[3], params[4], rest_bases, params[5], viewmat, projmat, bg);
        MTensor P[6] = {s.means, s.scales, s.quats, s.dc, s.rest, s.opac};
        MTensor A[6], B[6];
        for (int g = 0; g < 6; g++) {
            A[g] = upload(m1[g], P[g].shape(), DType::Float32);
            B[g] = upload(m2[g], P[g].shape(), DType::Float32);
        }
        MTensor gtt = upload(gt, {H, W, 3}, DType::Float32);
        MTensor win = gpu_zeros({11, 11}, DType::Float32);
        MTensor tv = upload(vis, {N}, DType::Float32), tg = upload(gn, {N}, DType::Float32),
                tm = upload(ms, {N}, DType::Float32);
        float ss[6], bb[6], cam[3] = {cam_pos[0], cam_pos[1], cam_pos[2]};
        for (int g = 0; g < 6; g++) { ss[g] = steps[g]; bb[g] = bc2s[g]; }
        auto [radii, loss] = msplat_train_step(
            N, s.means, s.scales, 1.0f, s.quats, s.vm, s.pm, fx, fy, cx, cy, (unsigned)H, (unsigned)W, tiles(W, H), clip,
            (unsigned)degree, (unsigned)dtu, cam, s.dc, s.rest, s.opac, s.bg, gtt, win, ssim_weight,
            1.0f / (float)(H * W * 3), rest_bases, 6, P, A, B, ss, bb, b1, b2, eps, tv, tg, tm,
            1.0f / (float)(H > W ? H : W), nullptr);
        msplat_gpu_sync();
        for (int g = 0; g < 6; g++) {
            download(params[g], P[g]);
            download(m1[g], A[g]);
            download(m2[g], B[g]);
        }
        download(vis, tv);
        download(gn, tg);
        download(ms, tm);
        download(radii_out, radii);
        if (loss_out) *loss_out = loss;
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

// PriorStep as plain host arrays (the layout of the oracle's Prior). The tensors the
// step updates (sky texture, gradient and moments, exposure rows and moments, exposure
// gradient, loss terms) are copied back.
struct RefPrior {
    int32_t aux, sky;
    float *sky_tex, *sky_grad, *sky_m1, *sky_m2;
    int32_t sky_w, sky_h;
    float sky_frame[9];
    float sky_step, sky_bc2;
    const float *prior_depth;
    const uint8_t *prior_aux;
    int32_t prior_w, prior_h, has_depth, has_sky_mask;
    float depth_weight, sky_weight, fill_weight, huber_delta, min_alpha;
    int32_t mask_photometric;
    float *loss_terms;
    int32_t exposure, cam_index, num_cams;
    float *expo_params, *expo_m1, *expo_m2, *expo_grad;
    float expo_step, expo_bc2, expo_reg;
    float log_max_scale_ratio;
};

namespace {
struct PriorTensors {
    PriorStep p;
    MTensor sky_tex, sky_grad, sky_m1, sky_m2, depth, aux, terms, expo, expo_m1, expo_m2, expo_grad;
    explicit PriorTensors(const RefPrior &r) {
        p.aux = r.aux != 0;
        p.sky = r.sky != 0;
        const std::vector<int64_t> tex = {r.sky_h, r.sky_w, 3};
        if (r.sky_tex) { sky_tex = upload(r.sky_tex, tex, DType::Float32); p.sky_tex = &sky_tex; }
        if (r.sky_grad) { sky_grad = upload(r.sky_grad, tex, DType::Float32); p.sky_grad = &sky_grad; }
        if (r.sky_m1) { sky_m1 = upload(r.sky_m1, tex, DType::Float32); p.sky_exp_avg = &sky_m1; }
        if (r.sky_m2) { sky_m2 = upload(r.sky_m2, tex, DType::Float32); p.sky_exp_avg_sq = &sky_m2; }
        p.sky_w = r.sky_w;
        p.sky_h = r.sky_h;
        std::memcpy(p.sky_frame, r.sky_frame, sizeof p.sky_frame);
        p.sky_step_size = r.sky_step;
        p.sky_bc2_sqrt = r.sky_bc2;
        if (r.prior_depth) { depth = upload(r.prior_depth, {r.prior_h, r.prior_w}, DType::Float32); p.prior_depth = &depth; }
        if (r.prior_aux) { aux = upload(r.prior_aux, {r.prior_h, r.prior_w, 4}, DType::UInt8); p.prior_aux = &aux; }
        p.prior_w = r.prior_w;
        p.prior_h = r.prior_h;
        p.has_depth = r.has_depth != 0;
        p.has_sky_mask = r.has_sky_mask != 0;
        p.depth_weight = r.depth_weight;
        p.sky_weight = r.sky_weight;
        p.fill_weight = r.fill_weight;
        p.huber_delta = r.huber_delta;
        p.min_alpha = r.min_alpha;
        p.mask_photometric = r.mask_photometric != 0;
        if (r.loss_terms) { terms = upload(r.loss_terms, {4}, DType::Float32); p.loss_terms = &terms; }
        p.exposure = r.exposure != 0;
        p.cam_index = r.cam_index;
        const std::vector<int64_t> ex = {r.num_cams, 12};
        if (r.expo_params) { expo = upload(r.expo_params, ex, DType::Float32); p.expo_params = &expo; }
        if (r.expo_m1) { expo_m1 = upload(r.expo_m1, ex, DType::Float32); p.expo_exp_avg = &expo_m1; }
        if (r.expo_m2) { expo_m2 = upload(r.expo_m2, ex, DType::Float32); p.expo_exp_avg_sq = &expo_m2; }
        if (r.expo_grad) { expo_grad = upload(r.expo_grad, {12}, DType::Float32); p.expo_grad = &expo_grad; }
        p.expo_step_size = r.expo_step;
        p.expo_bc2_sqrt = r.expo_bc2;
        p.expo_reg = r.expo_reg;
        p.log_max_scale_ratio = r.log_max_scale_ratio;
    }
    void back(const RefPrior &r) const {
        if (r.sky_tex) download(r.sky_tex, sky_tex);
        if (r.sky_grad) download(r.sky_grad, sky_grad);
        if (r.sky_m1) download(r.sky_m1, sky_m1);
        if (r.sky_m2) download(r.sky_m2, sky_m2);
        if (r.loss_terms) download(r.loss_terms, terms);
        if (r.expo_params) download(r.expo_params, expo);
        if (r.expo_m1) download(r.expo_m1, expo_m1);
        if (r.expo_m2) download(r.expo_m2, expo_m2);
        if (r.expo_grad) download(r.expo_grad, expo_grad);
    }
};
}  // namespace

// msplat_train_step with a PriorStep; arguments as ref_train_step, plus the prior and
// the aux outputs (expected-depth sum, final transmittance)
extern "C" int ref_train_step_prior(int N, float **params, int rest_bases, float **m1, float **m2,
                                    const float *viewmat, const float *projmat, float fx, float fy, float cx, float cy,
                                    int H, int W, float clip, int degree, int dtu, const float *cam_pos,
                                    const float *bg, const float *gt, float ssim_weight, const float *steps,
                                    const float *bc2s, float b1, float b2, float eps, float *vis, float *gn, float *ms,
                                    int32_t *radii_out, float *loss_out, const RefPrior *prior, float *depth_out,
                                    float *final_T_out) {
    try {
        Scene s(N, params[0], params[1], params[2], params[3], params[4], rest_bases, params[5], viewmat, projmat, bg);
        MTensor P[6] = {s.means, s.scales, s.quats, s.dc, s.rest, s.opac};
        MTensor A[6], B[6];
        for (int g = 0; g < 6; g++) {
            A[g] = upload(m1[g], P[g].shape(), DType::Float32);
            B[g] = upload(m2[g], P[g].shape(), DType::Float32);
        }
        MTensor gtt = upload(gt, {H, W, 3}, DType::Float32);
        MTensor win = gpu_zeros({11, 11}, DType::Float32);
        MTensor tv = upload(vis, {N}, DType::Float32), tg = upload(gn, {N}, DType::Float32),
                tm = upload(ms, {N}, DType::Float32);
        float ss[6], bb[6], cam[3] = {cam_pos[0], cam_pos[1], cam_pos[2]};
        for (int g = 0; g < 6; g++) { ss[g] = steps[g]; bb[g] = bc2s[g]; }
        PriorTensors pt(*prior);
        auto [radii, loss] = msplat_train_step(
            N, s.means, s.scales, 1.0f, s.quats, s.vm, s.pm, fx, fy, cx, cy, (unsigned)H, (unsigned)W, tiles(W, H), clip,
            (unsigned)degree, (unsigned)dtu, cam, s.dc, s.rest, s.opac, s.bg, gtt, win, ssim_weight,
            1.0f / (float)(H * W * 3), rest_bases, 6, P, A, B, ss, bb, b1, b2, eps, tv, tg, tm,
            1.0f / (float)(H > W ? H : W), &pt.p);
        msplat_gpu_sync();
        for (int g = 0; g < 6; g++) {
            download(params[g], P[g]);
            download(m1[g], A[g]);
            download(m2[g], B[g]);
        }
        download(vis, tv);
        download(gn, tg);
        download(ms, tm);
        download(radii_out, radii);
        pt.back(*prior);
        if (depth_out || final_T_out) {
            MTensor depth, fT;
            msplat_render_aux_outputs(depth, fT);
            if (depth_out && prior->aux) download(depth_out, depth);
            if (final_T_out) download(final_T_out, fT);
        }
        if (loss_out) *loss_out = loss;
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

// msplat_render with a PriorStep, and msplat_render_aux_outputs
extern "C" int ref_render_prior(int N, const float *means, const float *scales, const float *quats, const float *dc,
                                const float *rest, int rest_bases, const float *opac, const float *viewmat,
                                const float *projmat, float fx, float fy, float cx, float cy, int H, int W, float clip,
                                int degree, int dtu, const float *cam_pos, const float *bg, const RefPrior *prior,
                                float *out_img, float *depth_out, float *final_T_out) {
    try {
        Scene s(N, means, scales, quats, dc, rest, rest_bases, opac, viewmat, projmat, bg);
        float cam[3] = {cam_pos[0], cam_pos[1], cam_pos[2]};
        PriorTensors pt(*prior);
        MTensor img = msplat_render(N, s.means, s.scales, 1.0f, s.quats, s.vm, s.pm, fx, fy, cx, cy, (unsigned)H,
                                    (unsigned)W, tiles(W, H), clip, (unsigned)degree, (unsigned)dtu, cam, s.dc, s.rest,
                                    s.opac, s.bg, &pt.p);
        msplat_gpu_sync();
        download(out_img, img);
        MTensor depth, fT;
        msplat_render_aux_outputs(depth, fT);
        if (depth_out && prior->aux) download(depth_out, depth);
        if (final_T_out) download(final_T_out, fT);
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

extern "C" int ref_finalize_depth(const float *d, const float *t, uint32_t n, float inv_scale, float *od, float *oa) {
    try {
        MTensor dn = upload(d, {(int64_t)n}, DType::Float32), ft = upload(t, {(int64_t)n}, DType::Float32);
        MTensor o1 = gpu_zeros({(int64_t)n}, DType::Float32), o2 = gpu_zeros({(int64_t)n}, DType::Float32);
        msplat_finalize_depth(dn, ft, n, inv_scale, o1, o2);
        download(od, o1);
        download(oa, o2);
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

// msplat_densify over capacity-sized buffers: bufs[0..5] params, bufs[6..11] first
// moments, bufs[12..17] second moments, each with `cap` rows. Returns the new count.
extern "C" int ref_densify(int N, int cap, float grad_thr, float size_thr, float screen_thr, int check_screen,
                           float cull_alpha, float cull_scale, float cull_screen, int check_huge, const float *gn,
                           const float *vc, const float *ms, float half_max_dim, float **bufs, int fr_stride,
                           uint32_t seed) {
    try {
        const int64_t width[6] = {3, 3, 4, 3, fr_stride, 1};
        MTensor b[18];
        for (int i = 0; i < 18; i++) b[i] = upload(bufs[i], {cap, width[i % 6]}, DType::Float32);
        MTensor tg = upload(gn, {N}, DType::Float32), tv = upload(vc, {N}, DType::Float32),
                tm = upload(ms, {N}, DType::Float32);
        MTensor sc[8];
        for (auto &t : sc) t = gpu_zeros({cap}, DType::Int32);
        int out = msplat_densify(N, cap, grad_thr, size_thr, screen_thr, check_screen, cull_alpha, cull_scale,
                                 cull_screen, check_huge, tg, tv, tm, half_max_dim, b[0], b[1], b[2], b[3], b[4], b[5],
                                 fr_stride, b + 6, b + 12, sc[0], sc[1], sc[2], sc[3], sc[4], sc[5], sc[6], sc[7], seed);
        for (int i = 0; i < 18; i++) download(bufs[i], b[i]);
        return out;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

extern "C" void ref_opacity_reset(float *o, float *m1, float *m2, int n, float reset_logit) {
    MTensor a = upload(o, {n}, DType::Float32), b = upload(m1, {n}, DType::Float32), c = upload(m2, {n}, DType::Float32);
    msplat_opacity_reset(a, b, c, n, reset_logit);
    download(o, a);
    download(m1, b);
    download(m2, c);
}

extern "C" void ref_radix_sort(uint64_t *keys, int32_t *vals, uint32_t n, int key_bits) {
    int64_t m = n > 0 ? n : 1;
    MTensor k = upload(nullptr, {m}, DType::Int64), v = upload(nullptr, {m}, DType::Int32);
    std::memcpy(k.data_ptr(), keys, (size_t)n * 8);
    std::memcpy(v.data_ptr(), vals, (size_t)n * 4);
    msplat_radix_sort(k, v, n, key_bits);
    std::memcpy(keys, k.data_ptr(), (size_t)n * 8);
    std::memcpy(vals, v.data_ptr(), (size_t)n * 4);
}

extern "C" void ref_knn3(const float *pts, uint32_t n, float *out) {
    MTensor p = upload(pts, {(int64_t)n, 3}, DType::Float32), d = upload(nullptr, {(int64_t)n}, DType::Float32);
    msplat_knn3_mean_dist(p, n, d);
    download(out, d);
}

extern "C" void ref_init_gaussians(const float *means, const uint8_t *rgb, uint32_t n, uint32_t seed, float logit,
                                   float *scales, float *quats, float *dc, float *opac) {
    int64_t m = n;
    MTensor tm = upload(means, {m, 3}, DType::Float32), tc = upload(rgb, {m, 3}, DType::UInt8);
    MTensor s = upload(nullptr, {m, 3}, DType::Float32), q = upload(nullptr, {m, 4}, DType::Float32),
            d = upload(nullptr, {m, 3}, DType::Float32), o = upload(nullptr, {m, 1}, DType::Float32);
    msplat_init_gaussians(tm, tc, n, seed, logit, s, q, d, o);
    download(scales, s);
    download(quats, q);
    download(dc, d);
    download(opac, o);
}

extern "C" void ref_image_metrics(const float *a, const float *b, int h, int w, double *out) {
    MTensor ta = upload(a, {h, w, 3}, DType::Float32), tb = upload(b, {h, w, 3}, DType::Float32);
    msplat_image_metrics(ta, tb, h, w, out);
}

extern "C" void ref_pack_rgba8(const float *img, uint32_t n, uint8_t *out) {
    MTensor t = upload(img, {(int64_t)n, 3}, DType::Float32);
    msplat_pack_rgba8(t, n, out);
}

extern "C" void ref_resize_area(const void *src, int src_is_u8, int sw, int sh, float *out, int dw, int dh) {
    MTensor t = src_is_u8 ? upload(src, {sh, sw, 4}, DType::UInt8) : upload(src, {sh, sw, 3}, DType::Float32);
    MTensor d = upload(nullptr, {dh, dw, 3}, DType::Float32);
    msplat_resize_area(t, src_is_u8 != 0, sw, sh, d, dw, dh);
    msplat_gpu_sync();
    download(out, d);
}

extern "C" void ref_undistort(const float *src, int sw, int sh, const float *intr, const float *dist, int roi_x,
                              int roi_y, float *out, int dw, int dh) {
    MTensor t = upload(src, {sh, sw, 3}, DType::Float32);
    MTensor d = upload(nullptr, {dh, dw, 3}, DType::Float32);
    msplat_undistort(t, sw, sh, intr, dist, roi_x, roi_y, d, dw, dh);
    msplat_gpu_sync();
    download(out, d);
}

// saveGaussianPly / saveGaussianSplat / loadGaussianPly (core/src/loaders/save_gaussians.cpp)
// on host arrays: params[6] = means, scales, quats, dc, rest (n, rest_bases, 3), opacity.
extern "C" int ref_save_ply(int n, float **params, int rest_bases, const char *path, int step, float scale,
                            const float *translation, int keep_crs, int splat) {
    try {
        MTensor m = upload(params[0], {n, 3}, DType::Float32), s = upload(params[1], {n, 3}, DType::Float32),
                q = upload(params[2], {n, 4}, DType::Float32), d = upload(params[3], {n, 3}, DType::Float32),
                r = upload(params[4], {n, rest_bases, 3}, DType::Float32), o = upload(params[5], {n, 1}, DType::Float32);
        GaussianParams p{m, s, q, d, r, o, scale, {translation[0], translation[1], translation[2]}, keep_crs != 0};
        if (splat) saveGaussianSplat(path, p);
        else saveGaussianPly(path, p, step);
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}

extern "C" int ref_load_ply(const char *path, float scale, const float *translation, int keep_crs, int n,
                            int rest_bases, float **params, int *step) {
    try {
        LoadedGaussians g = loadGaussianPly(path, scale, translation, keep_crs != 0);
        if (g.means.size(0) != n || g.featuresRest.size(1) != rest_bases) return -2;
        download(params[0], g.means);
        download(params[1], g.scales);
        download(params[2], g.quats);
        download(params[3], g.featuresDc);
        download(params[4], g.featuresRest);
        download(params[5], g.opacities);
        *step = g.step;
        return 0;
    } catch (const std::exception &e) {
        return fail(e);
    }
}
