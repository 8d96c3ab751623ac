// core/hip/bindings_hip.cpp — msplat's backend interface (core/metal/bindings.h) on the
// HIP C ABI (include/msplat_hip.h). Dropped into the reference tree where core/metal
// sits, it lets Model, the loaders, python/bindings.cpp and tests/gradcheck.cpp run on
// an MI355X unchanged.
//
// Memory: the reference's MTensor on non-Apple builds is a host block that callers read
// and write through data_ptr() (core/include/metal_tensor.hpp:86-114), and results are
// visible there after msplat_gpu_sync (bindings.h:25). This adapter keeps a device mirror
// per host block (keyed by its base address):
//  - a tensor a call reads is uploaded before the launch, unless the device holds a
//    newer copy (a backend call wrote it and no sync has come since): then the device
//    copy is used as it is;
//  - a tensor a call writes is marked device-newer and stays on the device;
//  - msplat_gpu_sync downloads every device-newer tensor, after which host and device
//    agree and the host is authoritative again.
// So a training loop that syncs only when it reads (as Model and python/bindings.cpp
// do) keeps parameters, moments and statistics in HBM between syncs; per step only the
// read-only inputs (target image, matrices) cross PCIe. msplat_copy_buffer copies on the
// device when the source is device-newer. gpu_empty / gpu_zeros drop the mirror of a
// reused block address, and mirrors unused for many calls are freed once they exceed a
// byte budget.
//
// PriorStep's tensors follow the same rule: inputs are uploaded unless device-newer, and
// the ones a train step updates (sky texture, gradient and moments, exposure rows,
// moments and gradient, loss terms) become device-newer. The aux outputs of the last
// prior-guided frame are kept in two backend-owned tensors. GPU capture (a Metal tool)
// returns false.
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
    bool dev_newer = false;  // written on the device since the last sync
    uint64_t last_use = 0;
};

struct Backend {
    msplat_hip_ctx *ctx = nullptr;
    std::unordered_map<const void *, Mirror> mirrors;
    size_t mirror_bytes = 0;
    uint64_t clock = 0;
    MTensor out_img, radii, aux_depth, aux_T;
    bool aux_frame = false;
    bool timing = false;
    std::mutex mu;
    static constexpr size_t kBudget = size_t(16) << 30;  // idle mirrors are freed above this

    msplat_hip_ctx *get() {
        if (!ctx) check(msplat_hip_ctx_create(0, nullptr, &ctx), "context");
        return ctx;
    }
    static void check(int rc, const char *what) {
        if (rc != MSPLAT_OK) throw std::runtime_error(std::string("msplat HIP (") + what + "): " + msplat_hip_last_error());
    }
    // The block's mirror with at least `bytes`; a grown mirror keeps its device-newer bytes.
    Mirror &mirror(const MTensor &t, size_t bytes) {
        const void *key = t.data_ptr();
        Mirror &m = mirrors[key];
        m.last_use = ++clock;
        if (m.bytes < bytes) {
            void *d = nullptr;
            check(msplat_hip_malloc(bytes, 1, &d), "malloc");
            if (m.dev) {
                if (m.dev_newer) check(msplat_hip_memcpy(get(), d, m.dev, m.bytes, 2), "grow");
                check(msplat_hip_sync(get()), "sync");
                check(msplat_hip_free(m.dev), "free");
                mirror_bytes -= m.bytes;
            }
            if (m.dev_newer && bytes > m.bytes)  // the part the device never held comes from the host
                check(msplat_hip_memcpy(get(), (char *)d + m.bytes, (const char *)key + m.bytes, bytes - m.bytes, 0),
                      "upload");
            m.dev = d;
            m.bytes = bytes;
            mirror_bytes += bytes;
        }
        return m;
    }
    void *dev(const MTensor &t, size_t bytes) { return mirror(t, bytes).dev; }
    // An input: its device copy, uploaded unless the device copy is newer.
    float *up(const MTensor &t) {
        Mirror &m = mirror(t, t.nbytes());
        if (!m.dev_newer && t.nbytes()) check(msplat_hip_memcpy(get(), m.dev, t.data_ptr(), t.nbytes(), 0), "upload");
        return (float *)m.dev;
    }
    // Written on the device by the call just enqueued: host reads wait for msplat_gpu_sync.
    void wrote(const MTensor &t) {
        Mirror &m = mirror(t, t.nbytes());
        m.dev_newer = true;
    }
    void sync() {
        check(msplat_hip_sync(get()), "sync");
        for (auto &kv : mirrors)
            if (kv.second.dev_newer) {
                check(msplat_hip_memcpy(get(), (void *)kv.first, kv.second.dev, kv.second.bytes, 1), "download");
                kv.second.dev_newer = false;
            }
    }
    // A freshly allocated block at `key` has no mirror yet (a freed block's address reused).
    void forget(const void *key) {
        auto it = mirrors.find(key);
        if (it == mirrors.end()) return;
        if (it->second.dev) {
            check(msplat_hip_sync(get()), "sync");
            check(msplat_hip_free(it->second.dev), "free");
        }
        mirror_bytes -= it->second.bytes;
        mirrors.erase(it);
    }
    // Frees the least recently used clean mirrors while their total exceeds the budget.
    void trim() {
        while (mirror_bytes > kBudget) {
            auto victim = mirrors.end();
            for (auto it = mirrors.begin(); it != mirrors.end(); ++it)
                if (!it->second.dev_newer && (victim == mirrors.end() || it->second.last_use < victim->second.last_use))
                    victim = it;
            if (victim == mirrors.end()) return;
            forget(victim->first);
        }
    }
    void release() {
        for (auto &kv : mirrors)
            if (kv.second.dev) msplat_hip_free(kv.second.dev);
        mirrors.clear();
        mirror_bytes = 0;
        if (ctx) msplat_hip_ctx_destroy(ctx);
        ctx = nullptr;
        out_img = MTensor();
        radii = MTensor();
        aux_depth = MTensor();
        aux_T = MTensor();
        aux_frame = false;
    }
};

Backend g;

// A PriorStep on device mirrors; `written` lists the tensors a train step updates
struct DevicePrior {
    msplat_hip_prior p;
    std::vector<MTensor *> written;
};

DevicePrior device_prior(const PriorStep *ps, bool train) {
    DevicePrior d;
    std::memset(&d.p, 0, sizeof d.p);
    msplat_hip_prior &p = d.p;
    auto in = [](MTensor *t) -> float * { return t && t->numel() ? g.up(*t) : nullptr; };
    p.aux = ps->aux;
    p.sky = ps->sky;
    p.sky_tex = in(ps->sky_tex);
    p.sky_w = ps->sky_w;
    p.sky_h = ps->sky_h;
    std::memcpy(p.sky_frame, ps->sky_frame, sizeof p.sky_frame);
    if (train) {
        p.sky_grad = in(ps->sky_grad);
        p.sky_exp_avg = in(ps->sky_exp_avg);
        p.sky_exp_avg_sq = in(ps->sky_exp_avg_sq);
        p.sky_step_size = ps->sky_step_size;
        p.sky_bc2_sqrt = ps->sky_bc2_sqrt;
        p.prior_depth = in(ps->prior_depth);
        p.prior_aux = (const uint8_t *)in(ps->prior_aux);
        p.prior_w = ps->prior_w;
        p.prior_h = ps->prior_h;
        p.has_depth = ps->has_depth;
        p.has_sky_mask = ps->has_sky_mask;
        p.depth_weight = ps->depth_weight;
        p.sky_weight = ps->sky_weight;
        p.fill_weight = ps->fill_weight;
        p.huber_delta = ps->huber_delta;
        p.min_alpha = ps->min_alpha;
        p.mask_photometric = ps->mask_photometric;
        p.loss_terms = in(ps->loss_terms);
        p.exposure = ps->exposure;
        p.cam_index = ps->cam_index;
        if (ps->expo_params) p.num_cams = (int32_t)ps->expo_params->size(0);
        p.expo_params = in(ps->expo_params);
        p.expo_exp_avg = in(ps->expo_exp_avg);
        p.expo_exp_avg_sq = in(ps->expo_exp_avg_sq);
        p.expo_grad = in(ps->expo_grad);
        p.expo_step_size = ps->expo_step_size;
        p.expo_bc2_sqrt = ps->expo_bc2_sqrt;
        p.expo_reg = ps->expo_reg;
        p.log_max_scale_ratio = ps->log_max_scale_ratio;
        MTensor *upd[] = {ps->sky_tex, ps->sky_grad, ps->sky_exp_avg, ps->sky_exp_avg_sq, ps->loss_terms,
                          ps->expo_params, ps->expo_exp_avg, ps->expo_exp_avg_sq, ps->expo_grad};
        for (MTensor *t : upd)
            if (t && t->numel()) d.written.push_back(t);
    }
    return d;
}

// The aux outputs of the frame just rendered, copied into the backend's tensors
void keep_aux_outputs(unsigned H, unsigned W) {
    float *depth = nullptr, *fT = nullptr;
    Backend::check(msplat_hip_render_aux_outputs(g.get(), &depth, &fT), "aux outputs");
    MTensor *outs[2] = {&g.aux_depth, &g.aux_T};
    float *src[2] = {depth, fT};
    for (int i = 0; i < 2; i++) {
        MTensor &t = *outs[i];
        if (!t.defined() || t.size(0) != (int64_t)H || t.size(1) != (int64_t)W) {
            t = MTensor(MTensor::HostDevice{}, {(int64_t)H, (int64_t)W}, DType::Float32);
            g.forget(t.data_ptr());
        }
        Backend::check(msplat_hip_memcpy(g.get(), g.dev(t, t.nbytes()), src[i], t.nbytes(), 2), "aux copy");
        g.wrote(t);
    }
    g.aux_frame = true;
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

MTensor gpu_zeros(std::vector<int64_t> shape, DType dtype) {
    MTensor t(MTensor::HostDevice{}, std::move(shape), dtype);
    std::lock_guard<std::mutex> lk(g.mu);
    g.forget(t.data_ptr());
    return t;
}
MTensor gpu_empty(std::vector<int64_t> shape, DType dtype) { return gpu_zeros(std::move(shape), dtype); }

void msplat_commit() { Backend::check(msplat_hip_commit(g.get()), "commit"); }
void msplat_gpu_sync() {
    std::lock_guard<std::mutex> lk(g.mu);
    g.sync();
    g.trim();
}

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
    msplat_hip_gaussians gs = gaussians(num_points, means3d, scales, quats, features_dc, features_rest, opacities);
    msplat_hip_camera c = camera(viewmat, projmat, fx, fy, cx, cy, img_height, img_width, tile_bounds, clip_thresh,
                                 glob_scale, degree, degrees_to_use, cam_pos);
    float *img = nullptr;
    if (prior) {
        DevicePrior dp = device_prior(prior, false);
        Backend::check(msplat_hip_render_prior(g.get(), &gs, &c, background.data<float>(), &dp.p, &img), "render");
    } else {
        Backend::check(msplat_hip_render(g.get(), &gs, &c, background.data<float>(), &img), "render");
    }
    g.aux_frame = false;
    if (prior && prior->aux) keep_aux_outputs(img_height, img_width);
    if (!g.out_img.defined() || g.out_img.size(0) != (int64_t)img_height || g.out_img.size(1) != (int64_t)img_width) {
        g.out_img = MTensor(MTensor::HostDevice{}, {(int64_t)img_height, (int64_t)img_width, 3}, DType::Float32);
        g.forget(g.out_img.data_ptr());
    }
    Backend::check(msplat_hip_memcpy(g.get(), g.dev(g.out_img, g.out_img.nbytes()), img, g.out_img.nbytes(), 2), "image");
    g.wrote(g.out_img);
    return g.out_img;  // aliases the backend's cache until the next call; read after msplat_gpu_sync
}

void msplat_render_aux_outputs(MTensor &depth, MTensor &final_T) {
    std::lock_guard<std::mutex> lk(g.mu);
    if (!g.aux_frame) throw std::invalid_argument("msplat HIP backend: no prior-guided (aux) frame rendered");
    depth = g.aux_depth;  // alias the backend's tensors until the next frame; read after msplat_gpu_sync
    final_T = g.aux_T;
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
    if (!g.radii.defined() || g.radii.numel() != std::max(num_points, 1)) {
        g.radii = MTensor(MTensor::HostDevice{}, {std::max(num_points, 1)}, DType::Int32);
        g.forget(g.radii.data_ptr());
    }
    int32_t *d_radii = (int32_t *)g.dev(g.radii, g.radii.nbytes());
    float loss = 0.f;
    g.aux_frame = false;
    if (prior) {
        DevicePrior dp = device_prior(prior, true);
        Backend::check(msplat_hip_train_step_prior(g.get(), &gs, &c, background.data<float>(), g.up(gt), ssim_weight,
                                                   loss_inv_n, &a, &st, inv_max_dim, d_radii, &loss, &dp.p),
                       "train_step");
        for (MTensor *t : dp.written) g.wrote(*t);
        if (prior->aux) keep_aux_outputs(img_height, img_width);
    } else {
        Backend::check(msplat_hip_train_step(g.get(), &gs, &c, background.data<float>(), g.up(gt), ssim_weight,
                                             loss_inv_n, &a, &st, inv_max_dim, d_radii, &loss),
                       "train_step");
    }
    MTensor *written[] = {&means3d, &scales, &quats, &features_dc, &features_rest, &opacities,
                          &vis_counts, &xys_grad_norm, &max_2d_size, &g.radii};
    for (MTensor *t : written)
        if (t->numel()) g.wrote(*t);
    g.trim();
    for (int i = 0; i < 6; i++) {
        g.wrote(adam_exp_avg[i]);
        g.wrote(adam_exp_avg_sq[i]);
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
    for (int i = 0; i < 18; i++) g.wrote(*b[i]);
    return out;
}

void msplat_opacity_reset(MTensor &opacities, MTensor &exp_avg, MTensor &exp_avg_sq, int num_points,
                          float reset_logit) {
    std::lock_guard<std::mutex> lk(g.mu);
    float *o = g.up(opacities), *m1 = g.up(exp_avg), *m2 = g.up(exp_avg_sq);
    Backend::check(msplat_hip_opacity_reset(g.get(), o, m1, m2, num_points, reset_logit), "opacity_reset");
    g.wrote(opacities);
    g.wrote(exp_avg);
    g.wrote(exp_avg_sq);
}

void msplat_copy_buffer(MTensor &dst, const MTensor &src, size_t bytes) {
    if (!bytes) return;
    std::lock_guard<std::mutex> lk(g.mu);
    auto it = g.mirrors.find(src.data_ptr());
    if (it == g.mirrors.end() || !it->second.dev_newer) {  // the host holds the source: a plain copy
        if (g.mirrors.count(dst.data_ptr()) && g.mirrors[dst.data_ptr()].dev_newer) g.sync();
        std::memcpy(dst.data_ptr(), src.data_ptr(), bytes);
        return;
    }
    const void *s = g.mirror(src, bytes).dev;  // covers `bytes` (a view's mirror may hold fewer)
    float *d = g.up(dst);  // the destination's other bytes, then the source's prefix on the device
    s = g.mirrors.at(src.data_ptr()).dev;  // looked up again: up() may have rehashed the map
    Backend::check(msplat_hip_memcpy(g.get(), d, s, bytes, 2), "copy_buffer");
    g.wrote(dst);
}

void msplat_radix_sort(MTensor &keys, MTensor &vals, uint32_t n, int key_bits) {
    std::lock_guard<std::mutex> lk(g.mu);
    uint64_t *k = (uint64_t *)g.up(keys);
    int32_t *v = (int32_t *)g.up(vals);
    Backend::check(msplat_hip_radix_sort(g.get(), k, v, n, key_bits), "radix_sort");
    g.wrote(keys);
    g.wrote(vals);
}

void msplat_knn3_mean_dist(MTensor &points, uint32_t n, MTensor &mean_dist) {
    std::lock_guard<std::mutex> lk(g.mu);
    float *p = g.up(points);
    float *d = (float *)g.dev(mean_dist, mean_dist.nbytes());
    Backend::check(msplat_hip_knn3_mean_dist(g.get(), p, n, d), "knn3");
    g.wrote(mean_dist);
}

void msplat_init_gaussians(MTensor &means, MTensor &rgb, uint32_t n, uint32_t seed, float opacity_logit,
                           MTensor &scales, MTensor &quats, MTensor &features_dc, MTensor &opacities) {
    std::lock_guard<std::mutex> lk(g.mu);
    float *m = g.up(means);
    const uint8_t *c = (const uint8_t *)g.up(rgb);
    float *s = (float *)g.dev(scales, scales.nbytes()), *q = (float *)g.dev(quats, quats.nbytes());
    float *dc = (float *)g.dev(features_dc, features_dc.nbytes()), *o = (float *)g.dev(opacities, opacities.nbytes());
    Backend::check(msplat_hip_init_gaussians(g.get(), m, c, n, seed, opacity_logit, s, q, dc, o), "init");
    g.wrote(scales);
    g.wrote(quats);
    g.wrote(features_dc);
    g.wrote(opacities);
}

void msplat_resize_area(const MTensor &src, bool src_is_u8, int sw, int sh, MTensor &dst, int dw, int dh) {
    if (dw <= 0 || dh <= 0) return;
    std::lock_guard<std::mutex> lk(g.mu);
    const void *s = g.up(src);
    Backend::check(msplat_hip_resize_area(g.get(), s, src_is_u8 ? 1 : 0, sw, sh, (float *)g.dev(dst, dst.nbytes()), dw,
                                          dh),
                   "resize_area");
    g.wrote(dst);
}

void msplat_undistort(const MTensor &src, int sw, int sh, const float intr[4], const float dist[5], int roi_x,
                      int roi_y, MTensor &dst, int dw, int dh) {
    if (dw <= 0 || dh <= 0) return;
    std::lock_guard<std::mutex> lk(g.mu);
    const float *s = g.up(src);
    Backend::check(msplat_hip_undistort(g.get(), s, sw, sh, intr, dist, roi_x, roi_y,
                                        (float *)g.dev(dst, dst.nbytes()), dw, dh),
                   "undistort");
    g.wrote(dst);
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

void msplat_finalize_depth(const MTensor &depth_num, const MTensor &final_T, uint32_t n, float inv_scale,
                           MTensor &depth_out, MTensor &alpha_out) {
    std::lock_guard<std::mutex> lk(g.mu);
    float *d = g.up(depth_num), *t = g.up(final_T);
    float *od = (float *)g.dev(depth_out, depth_out.nbytes()), *oa = (float *)g.dev(alpha_out, alpha_out.nbytes());
    Backend::check(msplat_hip_finalize_depth(g.get(), d, t, n, inv_scale, od, oa), "finalize_depth");
    g.wrote(depth_out);
    g.wrote(alpha_out);
}
