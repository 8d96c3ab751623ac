// core/hip/bindings_hip.cpp — msplat's backend interface (core/metal/bindings.h) on the
// HIP C ABI (include/msplat_hip.h). Dropped into the reference tree where core/metal
// sits, it lets Model, the loaders, python/bindings.cpp and tests/gradcheck.cpp run on
// an MI355X unchanged.
//
// Memory: the reference's MTensor on non-Apple builds is a host block that callers read
// and write through data_ptr() (core/include/metal_tensor.hpp:86-114). This adapter keeps
// a device mirror per host block (keyed by its base address) and moves exactly the bytes
// each call reads and writes: inputs go up before the launch, written tensors come back
// after a sync. That keeps the reference's host-visible contract; the device-resident
// path (parameters and moments never leaving HBM) is the C ABI used directly, as
// msplat-street_amd/msplat/trainer.py does.
//
// Out of scope here (they throw): the prior-guided extras (PriorStep with anything
// enabled), the GT image pipeline (resize/undistort), aux depth outputs and GPU capture.
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "bindings.h"
#include "msplat_hip.h"

namespace {

struct Mirror {
    void *dev = nullptr;
    size_t bytes = 0;
};

struct Backend {
    msplat_hip_ctx *ctx = nullptr;
    std::unordered_map<const void *, Mirror> mirrors;
    MTensor out_img, radii;
    bool timing = false;
    std::mutex mu;

    msplat_hip_ctx *get() {
        if (!ctx) check(msplat_hip_ctx_create(0, nullptr, &ctx), "context");
        return ctx;
    }
    static void check(int rc, const char *what) {
        if (rc != MSPLAT_OK) throw std::runtime_error(std::string("msplat HIP (") + what + "): " + msplat_hip_last_error());
    }
    void *dev(const MTensor &t, size_t bytes) {
        const void *key = t.data_ptr();
        Mirror &m = mirrors[key];
        if (m.bytes < bytes) {
            if (m.dev) check(msplat_hip_free(m.dev), "free");
            check(msplat_hip_malloc(bytes, 1, &m.dev), "malloc");
            m.bytes = bytes;
        }
        return m.dev;
    }
    float *up(const MTensor &t) {
        void *d = dev(t, t.nbytes());
        check(msplat_hip_memcpy(get(), d, t.data_ptr(), t.nbytes(), 0), "upload");
        return (float *)d;
    }
    void down(MTensor &t) {
        void *d = dev(t, t.nbytes());
        check(msplat_hip_memcpy(get(), t.data_ptr(), d, t.nbytes(), 1), "download");
    }
    void release() {
        for (auto &kv : mirrors)
            if (kv.second.dev) msplat_hip_free(kv.second.dev);
        mirrors.clear();
        if (ctx) msplat_hip_ctx_destroy(ctx);
        ctx = nullptr;
        out_img = MTensor();
        radii = MTensor();
    }
};

Backend g;

void no_priors(const PriorStep *p) {
    if (p && (p->aux || p->sky || p->exposure || p->log_max_scale_ratio > 0.f))
        throw std::invalid_argument("msplat HIP backend: the prior-guided extras are not supported");
}

msplat_hip_camera camera(MTensor &viewmat, MTensor &projmat, float fx, float fy, float cx, float cy, unsigned H,
                         unsigned W, const std::tuple<int, int, int> &tiles, float clip, float glob, unsigned degree,
                         unsigned dtu, const float cam_pos[3]) {
    msplat_hip_camera c;
    std::memcpy(c.viewmat, viewmat.data_ptr(), sizeof c.viewmat);
    std::memcpy(c.projmat, projmat.data_ptr(), sizeof c.projmat);
    std::memcpy(c.cam_pos, cam_pos, sizeof c.cam_pos);
    c.fx = fx; c.fy = fy; c.cx = cx; c.cy = cy;
    c.width = W; c.height = H;
    c.tiles_x = std::get<0>(tiles);
    c.tiles_y = std::get<1>(tiles);
    c.clip_thresh = clip;
    c.glob_scale = glob;
    c.degree = degree;
    c.degrees_to_use = dtu;
    return c;
}

msplat_hip_gaussians gaussians(int n, MTensor &m, MTensor &s, MTensor &q, MTensor &dc, MTensor &rest, MTensor &o) {
    msplat_hip_gaussians gs;
    gs.num_points = n;
    gs.means = g.up(m);
    gs.scales = g.up(s);
    gs.quats = g.up(q);
    gs.features_dc = g.up(dc);
    gs.features_rest = rest.numel() ? g.up(rest) : gs.features_dc;
    gs.opacities = g.up(o);
    return gs;
}

}  // namespace

void cleanup_msplat_metal() {
    std::lock_guard<std::mutex> lk(g.mu);
    g.release();
}

void *msplat_device() { return g.get(); }

MTensor gpu_zeros(std::vector<int64_t> shape, DType dtype) { return MTensor(MTensor::HostDevice{}, std::move(shape), dtype); }
MTensor gpu_empty(std::vector<int64_t> shape, DType dtype) { return MTensor(MTensor::HostDevice{}, std::move(shape), dtype); }

void msplat_commit() { Backend::check(msplat_hip_commit(g.get()), "commit"); }
void msplat_gpu_sync() { Backend::check(msplat_hip_sync(g.get()), "sync"); }

bool msplat_begin_gpu_capture(const char *) {
    std::fprintf(stderr, "msplat: GPU capture is a Metal feature; use rocprofv3 on the HIP backend\n");
    return false;
}
void msplat_end_gpu_capture() {}

void msplat_enable_gpu_timing(bool enable) {
    g.timing = enable;
    Backend::check(msplat_hip_enable_stage_timing(g.get(), enable ? 1 : 0), "timing");
}

void msplat_drain_gpu_times(std::vector<double> &out) {
    out.clear();
    double ms[MSPLAT_NUM_STAGES];
    int64_t calls[MSPLAT_NUM_STAGES];
    int n = 0;
    Backend::check(msplat_hip_drain_stage_times(g.get(), ms, calls, MSPLAT_NUM_STAGES, &n), "drain");
    double total = 0;
    for (int i = 0; i < n; i++) total += ms[i];
    if (total > 0) out.push_back(total);
}

void msplat_drain_stage_times(std::vector<double> stage_times[], int max_stages, int &n_stages,
                              const char **stage_names) {
    double ms[MSPLAT_NUM_STAGES];
    int64_t calls[MSPLAT_NUM_STAGES];
    int n = 0;
    Backend::check(msplat_hip_drain_stage_times(g.get(), ms, calls, MSPLAT_NUM_STAGES, &n), "drain");
    n_stages = n < max_stages ? n : max_stages;
    for (int i = 0; i < n_stages; i++) {
        stage_times[i].clear();
        if (calls[i]) stage_times[i].push_back(ms[i]);
        if (stage_names) stage_names[i] = msplat_hip_stage_name(i);
    }
}

MTensor msplat_render(int num_points, MTensor &means3d, MTensor &scales, float glob_scale, MTensor &quats,
                      MTensor &viewmat, MTensor &projmat, float fx, float fy, float cx, float cy,
                      unsigned img_height, unsigned img_width, const std::tuple<int, int, int> tile_bounds,
                      float clip_thresh, unsigned degree, unsigned degrees_to_use, float cam_pos[3],
                      MTensor &features_dc, MTensor &features_rest, MTensor &opacities, MTensor &background,
                      const PriorStep *prior) {
    std::lock_guard<std::mutex> lk(g.mu);
    no_priors(prior);
    msplat_hip_gaussians gs = gaussians(num_points, means3d, scales, quats, features_dc, features_rest, opacities);
    msplat_hip_camera c = camera(viewmat, projmat, fx, fy, cx, cy, img_height, img_width, tile_bounds, clip_thresh,
                                 glob_scale, degree, degrees_to_use, cam_pos);
    float *img = nullptr;
    Backend::check(msplat_hip_render(g.get(), &gs, &c, background.data<float>(), &img), "render");
    if (!g.out_img.defined() || g.out_img.size(0) != (int64_t)img_height || g.out_img.size(1) != (int64_t)img_width)
        g.out_img = gpu_zeros({(int64_t)img_height, (int64_t)img_width, 3}, DType::Float32);
    Backend::check(msplat_hip_memcpy(g.get(), g.out_img.data_ptr(), img, g.out_img.nbytes(), 1), "image");
    return g.out_img;  // aliases the backend's cache until the next call (bindings.h:93-96)
}

void msplat_render_aux_outputs(MTensor &, MTensor &) {
    throw std::invalid_argument("msplat HIP backend: aux depth outputs belong to the prior-guided path");
}

std::tuple<MTensor, float> msplat_train_step(
    int num_points, MTensor &means3d, MTensor &scales, float glob_scale, MTensor &quats, MTensor &viewmat,
    MTensor &projmat, float fx, float fy, float cx, float cy, unsigned img_height, unsigned img_width,
    const std::tuple<int, int, int> tile_bounds, float clip_thresh, unsigned degree, unsigned degrees_to_use,
    float cam_pos[3], MTensor &features_dc, MTensor &features_rest, MTensor &opacities, MTensor &background,
    MTensor &gt, MTensor &window2d, float ssim_weight, float loss_inv_n, int features_rest_bases,
    int num_adam_groups, MTensor adam_params[], MTensor adam_exp_avg[], MTensor adam_exp_avg_sq[],
    float adam_step_sizes[], float adam_bc2_sqrts[], float adam_beta1, float adam_beta2, float adam_eps,
    MTensor &vis_counts, MTensor &xys_grad_norm, MTensor &max_2d_size, float inv_max_dim, const PriorStep *prior) {
    (void)window2d;
    (void)features_rest_bases;
    std::lock_guard<std::mutex> lk(g.mu);
    no_priors(prior);
    if (num_adam_groups < 6) throw std::invalid_argument("msplat HIP backend: six Adam groups expected");
    MTensor *params[6] = {&means3d, &scales, &quats, &features_dc, &features_rest, &opacities};
    for (int i = 0; i < 6; i++)
        if (adam_params[i].data_ptr() != params[i]->data_ptr())
            throw std::invalid_argument("msplat HIP backend: adam_params must alias the gaussian parameters");
    msplat_hip_gaussians gs = gaussians(num_points, means3d, scales, quats, features_dc, features_rest, opacities);
    msplat_hip_camera c = camera(viewmat, projmat, fx, fy, cx, cy, img_height, img_width, tile_bounds, clip_thresh,
                                 glob_scale, degree, degrees_to_use, cam_pos);
    msplat_hip_adam a;
    for (int i = 0; i < 6; i++) {
        a.exp_avg[i] = g.up(adam_exp_avg[i]);
        a.exp_avg_sq[i] = g.up(adam_exp_avg_sq[i]);
        a.step_size[i] = adam_step_sizes[i];
        a.bc2_sqrt[i] = adam_bc2_sqrts[i];
    }
    a.beta1 = adam_beta1;
    a.beta2 = adam_beta2;
    a.eps = adam_eps;
    msplat_hip_grad_stats st{g.up(vis_counts), g.up(xys_grad_norm), g.up(max_2d_size)};
    if (!g.radii.defined() || g.radii.numel() != num_points) g.radii = gpu_zeros({std::max(num_points, 1)}, DType::Int32);
    int32_t *d_radii = (int32_t *)g.dev(g.radii, g.radii.nbytes());
    float loss = 0.f;
    Backend::check(msplat_hip_train_step(g.get(), &gs, &c, background.data<float>(), g.up(gt), ssim_weight,
                                         loss_inv_n, &a, &st, inv_max_dim, d_radii, &loss),
                   "train_step");
    MTensor *written[] = {&means3d, &scales, &quats, &features_dc, &features_rest, &opacities,
                          &vis_counts, &xys_grad_norm, &max_2d_size, &g.radii};
    for (MTensor *t : written)
        if (t->numel()) g.down(*t);
    for (int i = 0; i < 6; i++) {
        g.down(adam_exp_avg[i]);
        g.down(adam_exp_avg_sq[i]);
    }
    return std::make_tuple(g.radii, loss);
}

int msplat_densify(int N, int buf_capacity, float grad_thresh, float size_thresh, float screen_thresh,
                   int check_screen, float cull_alpha_thresh, float cull_scale_thresh, float cull_screen_size,
                   int check_huge, MTensor &xys_grad_norm, MTensor &vis_counts, MTensor &max_2d_size,
                   float half_max_dim, MTensor &means_buf, MTensor &scales_buf, MTensor &quats_buf,
                   MTensor &featuresDc_buf, MTensor &featuresRest_buf, MTensor &opacities_buf, int fr_stride,
                   MTensor adam_exp_avg_buf[], MTensor adam_exp_avg_sq_buf[], MTensor &, MTensor &, MTensor &,
                   MTensor &, MTensor &, MTensor &, MTensor &, MTensor &, uint32_t seed) {
    std::lock_guard<std::mutex> lk(g.mu);
    MTensor *b[18] = {&means_buf, &scales_buf, &quats_buf, &featuresDc_buf, &featuresRest_buf, &opacities_buf};
    for (int i = 0; i < 6; i++) {
        b[6 + i] = &adam_exp_avg_buf[i];
        b[12 + i] = &adam_exp_avg_sq_buf[i];
    }
    float *d[18];
    for (int i = 0; i < 18; i++) d[i] = g.up(*b[i]);
    int32_t out = 0;
    Backend::check(msplat_hip_densify(g.get(), N, buf_capacity, grad_thresh, size_thresh, screen_thresh, check_screen,
                                      cull_alpha_thresh, cull_scale_thresh, cull_screen_size, check_huge,
                                      g.up(xys_grad_norm), g.up(vis_counts), g.up(max_2d_size), half_max_dim, d,
                                      fr_stride, seed, &out),
                   "densify");
    for (int i = 0; i < 18; i++) g.down(*b[i]);
    return out;
}

void msplat_opacity_reset(MTensor &opacities, MTensor &exp_avg, MTensor &exp_avg_sq, int num_points,
                          float reset_logit) {
    std::lock_guard<std::mutex> lk(g.mu);
    float *o = g.up(opacities), *m1 = g.up(exp_avg), *m2 = g.up(exp_avg_sq);
    Backend::check(msplat_hip_opacity_reset(g.get(), o, m1, m2, num_points, reset_logit), "opacity_reset");
    g.down(opacities);
    g.down(exp_avg);
    g.down(exp_avg_sq);
}

void msplat_copy_buffer(MTensor &dst, const MTensor &src, size_t bytes) {
    if (bytes) std::memcpy(dst.data_ptr(), src.data_ptr(), bytes);  // host blocks: a plain copy
}

void msplat_radix_sort(MTensor &keys, MTensor &vals, uint32_t n, int key_bits) {
    std::lock_guard<std::mutex> lk(g.mu);
    uint64_t *k = (uint64_t *)g.up(keys);
    int32_t *v = (int32_t *)g.up(vals);
    Backend::check(msplat_hip_radix_sort(g.get(), k, v, n, key_bits), "radix_sort");
    g.down(keys);
    g.down(vals);
}

void msplat_knn3_mean_dist(MTensor &points, uint32_t n, MTensor &mean_dist) {
    std::lock_guard<std::mutex> lk(g.mu);
    float *p = g.up(points);
    float *d = (float *)g.dev(mean_dist, mean_dist.nbytes());
    Backend::check(msplat_hip_knn3_mean_dist(g.get(), p, n, d), "knn3");
    g.down(mean_dist);
}

void msplat_init_gaussians(MTensor &means, MTensor &rgb, uint32_t n, uint32_t seed, float opacity_logit,
                           MTensor &scales, MTensor &quats, MTensor &features_dc, MTensor &opacities) {
    std::lock_guard<std::mutex> lk(g.mu);
    float *m = g.up(means);
    const uint8_t *c = (const uint8_t *)g.up(rgb);
    float *s = (float *)g.dev(scales, scales.nbytes()), *q = (float *)g.dev(quats, quats.nbytes());
    float *dc = (float *)g.dev(features_dc, features_dc.nbytes()), *o = (float *)g.dev(opacities, opacities.nbytes());
    Backend::check(msplat_hip_init_gaussians(g.get(), m, c, n, seed, opacity_logit, s, q, dc, o), "init");
    g.down(scales);
    g.down(quats);
    g.down(features_dc);
    g.down(opacities);
}

void msplat_resize_area(const MTensor &, bool, int, int, MTensor &, int, int) {
    throw std::invalid_argument("msplat HIP backend: the GT image pipeline (resize) is not part of this backend");
}

void msplat_undistort(const MTensor &, int, int, const float[4], const float[5], int, int, MTensor &, int, int) {
    throw std::invalid_argument("msplat HIP backend: the GT image pipeline (undistort) is not part of this backend");
}

void msplat_image_metrics(const MTensor &rendered, const MTensor &gt, int h, int w, double out[3]) {
    std::lock_guard<std::mutex> lk(g.mu);
    Backend::check(msplat_hip_image_metrics(g.get(), g.up(rendered), g.up(gt), h, w, out), "metrics");
}

void msplat_pack_rgba8(const MTensor &img, uint32_t n, uint8_t *out) {
    std::lock_guard<std::mutex> lk(g.mu);
    float *d = g.up(img);
    void *o = nullptr;
    Backend::check(msplat_hip_malloc((size_t)n * 4, 0, &o), "malloc");
    int rc = msplat_hip_pack_rgba8(g.get(), d, n, (uint8_t *)o);
    if (rc == MSPLAT_OK) rc = msplat_hip_memcpy(g.get(), out, o, (size_t)n * 4, 1);
    msplat_hip_free(o);
    Backend::check(rc, "pack_rgba8");
}

void msplat_finalize_depth(const MTensor &, const MTensor &, uint32_t, float, MTensor &, MTensor &) {
    throw std::invalid_argument("msplat HIP backend: expected depth belongs to the prior-guided path");
}
