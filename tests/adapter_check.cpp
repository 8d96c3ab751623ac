// Runs the drop-in C++ backend (msplat-street_amd/csrc/host/bindings_hip.cpp, which
// implements the reference's core/metal/bindings.h on the HIP C ABI) the way the
// reference's callers use it: MTensors from gpu_empty / gpu_zeros filled through
// data_ptr(), views of capacity buffers, train steps without a sync in between, a sync
// before host reads, host writes after a sync, msplat_copy_buffer into a grown buffer,
// densify and opacity reset on the capacity buffers, the image pipeline. Every result is
// compared, bit for bit, with the same calls made on the C ABI directly with device
// buffers. Built by oracle/Makefile (target adapter) against the reference's headers;
// prints one line per check and exits 0 when all pass.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <random>
#include <vector>

#include "bindings.h"
#include "msplat_hip.h"

namespace {

int failures = 0;

void expect(bool ok, const char *what) {
    std::printf("%-58s %s\n", what, ok ? "ok" : "MISMATCH");
    if (!ok) failures++;
}

void rc(int code, const char *what) {
    if (code != MSPLAT_OK) {
        std::printf("C ABI call failed (%s): %s\n", what, msplat_hip_last_error());
        std::exit(2);
    }
}

// device buffer holding a copy of host data
struct Dev {
    void *p = nullptr;
    size_t bytes = 0;
    Dev(const void *src, size_t b, msplat_hip_ctx *ctx) : bytes(b) {
        rc(msplat_hip_malloc(b ? b : 4, 1, &p), "malloc");
        if (b && src) rc(msplat_hip_memcpy(ctx, p, src, b, 0), "upload");
    }
    ~Dev() { msplat_hip_free(p); }
    float *f() const { return (float *)p; }
    std::vector<float> get(msplat_hip_ctx *ctx, size_t n) const {
        std::vector<float> out(n);
        rc(msplat_hip_memcpy(ctx, out.data(), p, 4 * n, 1), "download");
        return out;
    }
};

bool same(const float *a, const float *b, size_t n) { return std::memcmp(a, b, 4 * n) == 0; }

}  // namespace

int run() {
    std::mt19937 rng(7);
    std::normal_distribution<float> N01(0.f, 1.f);
    std::uniform_real_distribution<float> U(-1.f, 1.f);
    const int n = 3000, cap = 4 * n, W = 97, H = 71, RB = 15;  // SH degree 3
    const int widths[6] = {3, 3, 4, 3, 3 * RB, 1};

    // capacity buffers and their views, filled on the host as Model::setupOptimizers does
    MTensor bufs[6], m1[6], m2[6];
    for (int k = 0; k < 6; k++) {
        std::vector<int64_t> shape = k == 4 ? std::vector<int64_t>{cap, RB, 3} : std::vector<int64_t>{cap, widths[k]};
        bufs[k] = gpu_zeros(shape, DType::Float32);
        m1[k] = gpu_zeros(shape, DType::Float32);
        m2[k] = gpu_zeros(shape, DType::Float32);
        float *p = bufs[k].data<float>();
        for (int i = 0; i < n * widths[k]; i++) {
            float v = 0.f;
            switch (k) {
            case 0: v = 1.2f * U(rng); break;
            case 1: v = -3.2f + 0.4f * N01(rng); break;
            case 2: v = N01(rng); break;
            case 3: v = 0.5f * N01(rng); break;
            case 4: v = 0.1f * N01(rng); break;
            default: v = N01(rng);
            }
            p[i] = v;
        }
    }
    MTensor view[6], e1[6], e2[6];
    for (int k = 0; k < 6; k++) {
        view[k] = bufs[k].view(n);
        e1[k] = m1[k].view(n);
        e2[k] = m2[k].view(n);
    }
    // camera at (2.6, 0.7, 2.9) looking at the origin: world -> camera (OpenCV axes)
    const float eye[3] = {2.6f, 0.7f, 2.9f};
    float f[3] = {-eye[0], -eye[1], -eye[2]};
    const float fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (float &v : f) v /= fl;
    float r[3] = {f[1] * 0.f - f[2] * 1.f, f[2] * 0.f - f[0] * 0.f, f[0] * 1.f - f[1] * 0.f};
    const float rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (float &v : r) v /= rl;
    const float u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    const float R[9] = {r[0], r[1], r[2], -u[0], -u[1], -u[2], f[0], f[1], f[2]};
    float vm[16] = {}, pm[16] = {};
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) vm[4 * i + j] = R[3 * i + j];
        vm[4 * i + 3] = -(R[3 * i] * eye[0] + R[3 * i + 1] * eye[1] + R[3 * i + 2] * eye[2]);
    }
    vm[15] = 1.f;
    const float fx = 90.f, fy = 88.f, cx = 48.3f, cy = 35.6f;
    const float P[16] = {2 * fx / W, 0, 0, 0, 0, 2 * fy / H, 0, 0, 0, 0, 1.000002f, -0.002000002f, 0, 0, 1, 0};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 4; k++) pm[4 * i + j] += P[4 * i + k] * vm[4 * k + j];
    MTensor viewmat = gpu_empty({4, 4}, DType::Float32), projmat = gpu_empty({4, 4}, DType::Float32);
    std::memcpy(viewmat.data_ptr(), vm, sizeof vm);
    std::memcpy(projmat.data_ptr(), pm, sizeof pm);
    float cam_pos[3] = {eye[0], eye[1], eye[2]};
    MTensor bg = gpu_empty({3}, DType::Float32), gt = gpu_empty({H, W, 3}, DType::Float32),
            win = gpu_zeros({11, 11}, DType::Float32);
    const float bgv[3] = {0.61f, 0.01f, 0.4f};
    std::memcpy(bg.data_ptr(), bgv, sizeof bgv);
    for (int i = 0; i < H * W * 3; i++) gt.data<float>()[i] = 0.5f + 0.4f * U(rng);
    const auto tiles = std::make_tuple((W + 15) / 16, (H + 15) / 16, 1);
    MTensor vis = gpu_zeros({cap}, DType::Float32), gn = gpu_zeros({cap}, DType::Float32),
            ms = gpu_zeros({cap}, DType::Float32);
    MTensor visv = vis.view(n), gnv = gn.view(n), msv = ms.view(n);

    // the same state on the C ABI
    msplat_hip_ctx *ctx = nullptr;
    rc(msplat_hip_ctx_create(0, nullptr, &ctx), "ctx");
    std::vector<Dev *> dp, d1, d2;
    for (int k = 0; k < 6; k++) {
        dp.push_back(new Dev(bufs[k].data_ptr(), bufs[k].nbytes(), ctx));
        d1.push_back(new Dev(m1[k].data_ptr(), m1[k].nbytes(), ctx));
        d2.push_back(new Dev(m2[k].data_ptr(), m2[k].nbytes(), ctx));
    }
    Dev dvis(vis.data_ptr(), vis.nbytes(), ctx), dgn(gn.data_ptr(), gn.nbytes(), ctx), dms(ms.data_ptr(), ms.nbytes(), ctx);
    Dev dgt(gt.data_ptr(), gt.nbytes(), ctx);
    msplat_hip_camera c{};
    std::memcpy(c.viewmat, vm, sizeof vm);
    std::memcpy(c.projmat, pm, sizeof pm);
    std::memcpy(c.cam_pos, cam_pos, sizeof cam_pos);
    c.fx = fx; c.fy = fy; c.cx = cx; c.cy = cy;
    c.width = W; c.height = H;
    c.tiles_x = std::get<0>(tiles); c.tiles_y = std::get<1>(tiles);
    c.clip_thresh = 0.01f; c.glob_scale = 1.f; c.degree = 3; c.degrees_to_use = 3;
    auto gauss = [&](int count) {
        msplat_hip_gaussians g{count, dp[0]->f(), dp[1]->f(), dp[2]->f(), dp[3]->f(), dp[4]->f(), dp[5]->f()};
        return g;
    };

    // 1. render
    MTensor img = msplat_render(n, view[0], view[1], 1.f, view[2], viewmat, projmat, fx, fy, cx, cy, H, W, tiles, 0.01f,
                                3, 3, cam_pos, view[3], view[4], view[5], bg, nullptr);
    msplat_gpu_sync();
    float *dimg = nullptr;
    msplat_hip_gaussians g0 = gauss(n);
    rc(msplat_hip_render(ctx, &g0, &c, bgv, &dimg), "render");
    std::vector<float> ref(H * W * 3);
    rc(msplat_hip_memcpy(ctx, ref.data(), dimg, 4 * ref.size(), 1), "image");
    expect(same(img.data<float>(), ref.data(), ref.size()), "render: image after msplat_gpu_sync");

    // 2. three train steps without a sync in between, then one sync
    float ss[6], bc2[6];
    const float lrs[6] = {1.6e-4f, 5e-3f, 1e-3f, 2.5e-3f, 1.25e-4f, 5e-2f};
    float losses[3], ref_losses[3];
    for (int t = 1; t <= 3; t++) {
        const float b1 = 1.f - (float)std::pow(0.9, t), b2 = 1.f - (float)std::pow(0.999, t);
        for (int k = 0; k < 6; k++) {
            ss[k] = lrs[k] / b1;
            bc2[k] = std::sqrt(b2);
        }
        auto res = msplat_train_step(n, view[0], view[1], 1.f, view[2], viewmat, projmat, fx, fy, cx, cy, H, W, tiles,
                                     0.01f, 3, 3, cam_pos, view[3], view[4], view[5], bg, gt, win, 0.2f,
                                     1.f / (float)(H * W * 3), RB, 6, view, e1, e2, ss, bc2, 0.9f, 0.999f, 1e-8f, visv,
                                     gnv, msv, 1.f / (float)W, nullptr);
        losses[t - 1] = std::get<1>(res);
        msplat_hip_adam a{};
        for (int k = 0; k < 6; k++) {
            a.exp_avg[k] = d1[k]->f();
            a.exp_avg_sq[k] = d2[k]->f();
            a.step_size[k] = ss[k];
            a.bc2_sqrt[k] = bc2[k];
        }
        a.beta1 = 0.9f; a.beta2 = 0.999f; a.eps = 1e-8f;
        msplat_hip_grad_stats st{dvis.f(), dgn.f(), dms.f()};
        msplat_hip_gaussians g = gauss(n);
        rc(msplat_hip_train_step(ctx, &g, &c, bgv, dgt.f(), 0.2f, 1.f / (float)(H * W * 3), &a, &st, 1.f / (float)W,
                                 nullptr, &ref_losses[t - 1]),
           "train_step");
    }
    msplat_gpu_sync();
    bool ok = std::memcmp(losses, ref_losses, sizeof losses) == 0;
    for (int k = 0; k < 6; k++) {
        ok = ok && same(bufs[k].data<float>(), dp[k]->get(ctx, (size_t)n * widths[k]).data(), (size_t)n * widths[k]);
        ok = ok && same(m1[k].data<float>(), d1[k]->get(ctx, (size_t)n * widths[k]).data(), (size_t)n * widths[k]);
        ok = ok && same(m2[k].data<float>(), d2[k]->get(ctx, (size_t)n * widths[k]).data(), (size_t)n * widths[k]);
    }
    ok = ok && same(vis.data<float>(), dvis.get(ctx, n).data(), n) && same(gn.data<float>(), dgn.get(ctx, n).data(), n);
    expect(ok, "train_step x3: losses, parameters, moments, statistics");

    // 3. a host write after the sync is seen by the next call (the gradcheck pattern)
    bufs[0].data<float>()[5] += 0.01f;
    bufs[5].data<float>()[9] -= 0.3f;
    rc(msplat_hip_memcpy(ctx, dp[0]->f(), bufs[0].data_ptr(), 4 * (size_t)n * 3, 0), "upload");
    rc(msplat_hip_memcpy(ctx, dp[5]->f(), bufs[5].data_ptr(), 4 * (size_t)n, 0), "upload");
    img = msplat_render(n, view[0], view[1], 1.f, view[2], viewmat, projmat, fx, fy, cx, cy, H, W, tiles, 0.01f, 3, 3,
                        cam_pos, view[3], view[4], view[5], bg, nullptr);
    msplat_gpu_sync();
    g0 = gauss(n);
    rc(msplat_hip_render(ctx, &g0, &c, bgv, &dimg), "render");
    rc(msplat_hip_memcpy(ctx, ref.data(), dimg, 4 * ref.size(), 1), "image");
    expect(same(img.data<float>(), ref.data(), ref.size()), "render after host writes to parameters");

    // 4. grow the means buffer on the device (Model::ensureCapacity), densify, reset
    MTensor grown = gpu_zeros({2 * cap, 3}, DType::Float32);
    msplat_train_step(n, view[0], view[1], 1.f, view[2], viewmat, projmat, fx, fy, cx, cy, H, W, tiles, 0.01f, 3, 3,
                      cam_pos, view[3], view[4], view[5], bg, gt, win, 0.2f, 1.f / (float)(H * W * 3), RB, 6, view, e1,
                      e2, ss, bc2, 0.9f, 0.999f, 1e-8f, visv, gnv, msv, 1.f / (float)W, nullptr);
    {
        msplat_hip_adam a{};
        for (int k = 0; k < 6; k++) {
            a.exp_avg[k] = d1[k]->f();
            a.exp_avg_sq[k] = d2[k]->f();
            a.step_size[k] = ss[k];
            a.bc2_sqrt[k] = bc2[k];
        }
        a.beta1 = 0.9f; a.beta2 = 0.999f; a.eps = 1e-8f;
        msplat_hip_grad_stats st{dvis.f(), dgn.f(), dms.f()};
        msplat_hip_gaussians g = gauss(n);
        rc(msplat_hip_train_step(ctx, &g, &c, bgv, dgt.f(), 0.2f, 1.f / (float)(H * W * 3), &a, &st, 1.f / (float)W,
                                 nullptr, nullptr),
           "train_step");
    }
    msplat_copy_buffer(grown, bufs[0], bufs[0].nbytes());  // device-newer source: copied on the device
    msplat_gpu_sync();
    expect(same(grown.data<float>(), dp[0]->get(ctx, (size_t)cap * 3).data(), (size_t)cap * 3),
           "copy_buffer of a device-newer buffer");
    MTensor scratch[8];
    for (auto &t : scratch) t = gpu_zeros({cap}, DType::Int32);
    const uint32_t seed = 600u * 2654435761u;
    int nn = msplat_densify(n, cap, 2e-5f, 0.05f, 0.05f, 1, 0.1f, 0.5f, 0.15f, 1, gn, vis, ms, 48.5f, bufs[0], bufs[1],
                            bufs[2], bufs[3], bufs[4], bufs[5], 3 * RB, m1, m2, scratch[0], scratch[1], scratch[2],
                            scratch[3], scratch[4], scratch[5], scratch[6], scratch[7], seed);
    msplat_opacity_reset(bufs[5], m1[5], m2[5], nn, -1.3862944f);
    msplat_gpu_sync();
    float *d18[18];
    for (int k = 0; k < 6; k++) {
        d18[k] = dp[k]->f();
        d18[6 + k] = d1[k]->f();
        d18[12 + k] = d2[k]->f();
    }
    int32_t ref_n = 0;
    rc(msplat_hip_densify(ctx, n, cap, 2e-5f, 0.05f, 0.05f, 1, 0.1f, 0.5f, 0.15f, 1, dgn.f(), dvis.f(), dms.f(), 48.5f,
                          d18, 3 * RB, seed, &ref_n),
       "densify");
    rc(msplat_hip_opacity_reset(ctx, dp[5]->f(), d1[5]->f(), d2[5]->f(), ref_n, -1.3862944f), "reset");
    ok = nn == ref_n && nn != n;
    for (int k = 0; k < 6; k++) {
        ok = ok && same(bufs[k].data<float>(), dp[k]->get(ctx, (size_t)nn * widths[k]).data(), (size_t)nn * widths[k]);
        ok = ok && same(m1[k].data<float>(), d1[k]->get(ctx, (size_t)nn * widths[k]).data(), (size_t)nn * widths[k]);
    }
    std::printf("densify: %d -> %d gaussians\n", n, nn);
    expect(ok, "densify + opacity reset on the capacity buffers");

    // 5. image pipeline
    MTensor rgba = gpu_empty({60, 80, 4}, DType::UInt8), lvl = gpu_empty({25, 33, 3}, DType::Float32),
            und = gpu_empty({20, 30, 3}, DType::Float32);
    for (int i = 0; i < 60 * 80 * 4; i++) rgba.data<uint8_t>()[i] = (uint8_t)(rng() & 255);
    const float intr[4] = {30.f, 31.f, 16.2f, 12.9f}, dist[5] = {-0.2f, 0.05f, 0.001f, -0.002f, 0.01f};
    msplat_resize_area(rgba, true, 80, 60, lvl, 33, 25);
    msplat_undistort(lvl, 33, 25, intr, dist, 1, 2, und, 30, 20);
    msplat_gpu_sync();
    Dev drgba(rgba.data_ptr(), rgba.nbytes(), ctx);
    Dev dlvl(lvl.data_ptr(), lvl.nbytes(), ctx), dund(und.data_ptr(), und.nbytes(), ctx);
    rc(msplat_hip_resize_area(ctx, drgba.p, 1, 80, 60, dlvl.f(), 33, 25), "resize");
    rc(msplat_hip_undistort(ctx, dlvl.f(), 33, 25, intr, dist, 1, 2, dund.f(), 30, 20), "undistort");
    expect(same(lvl.data<float>(), dlvl.get(ctx, 25 * 33 * 3).data(), 25 * 33 * 3) &&
               same(und.data<float>(), dund.get(ctx, 20 * 30 * 3).data(), 20 * 30 * 3),
           "resize_area + undistort");

    // 6. prior-guided steps (aux rasterizer, learned sky, depth and sky priors, exposure,
    // needle cap) without a sync in between, a prior render, its aux outputs, finalize_depth
    {
        for (int k = 0; k < 6; k++) {  // views of the densified count
            view[k] = bufs[k].view(nn);
            e1[k] = m1[k].view(nn);
            e2[k] = m2[k].view(nn);
        }
        visv = vis.view(nn);
        gnv = gn.view(nn);
        msv = ms.view(nn);
        const int pw = W / 2 + 3, ph = H / 2 - 1, sw = 32, sh = 16;
        MTensor pdepth = gpu_zeros({ph, pw}, DType::Float32), paux = gpu_zeros({ph, pw, 4}, DType::UInt8);
        for (int y = 0; y < ph; y++)
            for (int x = 0; x < pw; x++) {
                float dv = 3.2f * (1.f + 0.25f * std::sin(0.3f * x) * std::cos(0.2f * y));
                if ((x * 7 + y * 3) % 23 == 0) dv = 0.f;
                pdepth.data<float>()[y * pw + x] = dv;
                uint8_t *a = paux.data<uint8_t>() + 4 * (y * pw + x);
                a[0] = (uint8_t)(60 + (x * 13 + y * 7) % 190);
                a[1] = y < ph / 4 ? 255 : 0;
                a[2] = (x + 2 * y) % 11 == 0 ? 0 : 255;
            }
        MTensor sky = gpu_zeros({sh, sw, 3}, DType::Float32), skyg = gpu_zeros({sh, sw, 3}, DType::Float32),
                skym1 = gpu_zeros({sh, sw, 3}, DType::Float32), skym2 = gpu_zeros({sh, sw, 3}, DType::Float32);
        for (int i = 0; i < sw * sh * 3; i++) sky.data<float>()[i] = 0.5f + 0.3f * U(rng);
        MTensor expo = gpu_zeros({2, 12}, DType::Float32), em1 = gpu_zeros({2, 12}, DType::Float32),
                em2 = gpu_zeros({2, 12}, DType::Float32), egrad = gpu_zeros({12}, DType::Float32),
                terms = gpu_zeros({4}, DType::Float32);
        for (int k = 0; k < 24; k++) expo.data<float>()[k] = ((k % 12) % 4 == 0 && k % 12 < 9 ? 1.f : 0.f) + 0.05f * U(rng);
        PriorStep pr;
        pr.aux = true;
        pr.sky = true;
        pr.sky_tex = &sky; pr.sky_grad = &skyg; pr.sky_exp_avg = &skym1; pr.sky_exp_avg_sq = &skym2;
        pr.sky_w = sw; pr.sky_h = sh;
        const float frame[9] = {0, -1, 0, 0, 0, 1, 1, 0, 0};
        std::memcpy(pr.sky_frame, frame, sizeof frame);
        pr.sky_step_size = 0.01f; pr.sky_bc2_sqrt = 0.0316f;
        pr.prior_depth = &pdepth; pr.prior_aux = &paux; pr.prior_w = pw; pr.prior_h = ph;
        pr.has_depth = pr.has_sky_mask = true;
        pr.depth_weight = 0.5f; pr.sky_weight = 0.3f; pr.fill_weight = 0.1f;
        pr.loss_terms = &terms;
        pr.exposure = true; pr.cam_index = 1;
        pr.expo_params = &expo; pr.expo_exp_avg = &em1; pr.expo_exp_avg_sq = &em2; pr.expo_grad = &egrad;
        pr.expo_step_size = 0.002f; pr.expo_bc2_sqrt = 0.0316f; pr.expo_reg = 0.01f;
        pr.log_max_scale_ratio = std::log(3.f);
        // the same prior on the C ABI
        Dev dpd(pdepth.data_ptr(), pdepth.nbytes(), ctx), dpa(paux.data_ptr(), paux.nbytes(), ctx);
        Dev dsky(sky.data_ptr(), sky.nbytes(), ctx), dskyg(skyg.data_ptr(), skyg.nbytes(), ctx),
            dskym1(skym1.data_ptr(), skym1.nbytes(), ctx), dskym2(skym2.data_ptr(), skym2.nbytes(), ctx);
        Dev dexpo(expo.data_ptr(), expo.nbytes(), ctx), dem1(em1.data_ptr(), em1.nbytes(), ctx),
            dem2(em2.data_ptr(), em2.nbytes(), ctx), degrad(egrad.data_ptr(), egrad.nbytes(), ctx),
            dterms(terms.data_ptr(), terms.nbytes(), ctx);
        msplat_hip_prior hp{};
        hp.aux = 1; hp.sky = 1;
        hp.sky_tex = dsky.f(); hp.sky_grad = dskyg.f(); hp.sky_exp_avg = dskym1.f(); hp.sky_exp_avg_sq = dskym2.f();
        hp.sky_w = sw; hp.sky_h = sh;
        std::memcpy(hp.sky_frame, frame, sizeof frame);
        hp.sky_step_size = 0.01f; hp.sky_bc2_sqrt = 0.0316f;
        hp.prior_depth = dpd.f(); hp.prior_aux = (const uint8_t *)dpa.p; hp.prior_w = pw; hp.prior_h = ph;
        hp.has_depth = hp.has_sky_mask = 1;
        hp.depth_weight = 0.5f; hp.sky_weight = 0.3f; hp.fill_weight = 0.1f;
        hp.huber_delta = pr.huber_delta; hp.min_alpha = pr.min_alpha; hp.mask_photometric = 1;
        hp.loss_terms = dterms.f();
        hp.exposure = 1; hp.cam_index = 1; hp.num_cams = 2;
        hp.expo_params = dexpo.f(); hp.expo_exp_avg = dem1.f(); hp.expo_exp_avg_sq = dem2.f(); hp.expo_grad = degrad.f();
        hp.expo_step_size = 0.002f; hp.expo_bc2_sqrt = 0.0316f; hp.expo_reg = 0.01f;
        hp.log_max_scale_ratio = std::log(3.f);
        float pl[2], rpl[2];
        for (int t = 0; t < 2; t++) {
            auto res = msplat_train_step(nn, view[0], view[1], 1.f, view[2], viewmat, projmat, fx, fy, cx, cy, H, W,
                                         tiles, 0.01f, 3, 3, cam_pos, view[3], view[4], view[5], bg, gt, win, 0.2f,
                                         1.f / (float)(H * W * 3), RB, 6, view, e1, e2, ss, bc2, 0.9f, 0.999f, 1e-8f,
                                         visv, gnv, msv, 1.f / (float)W, &pr);
            pl[t] = std::get<1>(res);
            msplat_hip_adam a{};
            for (int k = 0; k < 6; k++) {
                a.exp_avg[k] = d1[k]->f();
                a.exp_avg_sq[k] = d2[k]->f();
                a.step_size[k] = ss[k];
                a.bc2_sqrt[k] = bc2[k];
            }
            a.beta1 = 0.9f; a.beta2 = 0.999f; a.eps = 1e-8f;
            msplat_hip_grad_stats st{dvis.f(), dgn.f(), dms.f()};
            msplat_hip_gaussians gg = gauss(nn);
            rc(msplat_hip_train_step_prior(ctx, &gg, &c, bgv, dgt.f(), 0.2f, 1.f / (float)(H * W * 3), &a, &st,
                                           1.f / (float)W, nullptr, &rpl[t], &hp),
               "train_step_prior");
        }
        MTensor pimg = msplat_render(nn, view[0], view[1], 1.f, view[2], viewmat, projmat, fx, fy, cx, cy, H, W, tiles,
                                     0.01f, 3, 3, cam_pos, view[3], view[4], view[5], bg, &pr);
        MTensor dnum, fT;
        msplat_render_aux_outputs(dnum, fT);
        MTensor depth_out = gpu_zeros({H, W}, DType::Float32), alpha_out = gpu_zeros({H, W}, DType::Float32);
        msplat_finalize_depth(dnum, fT, (uint32_t)(H * W), 0.5f, depth_out, alpha_out);
        msplat_gpu_sync();
        float *rimg = nullptr, *rdep = nullptr, *rT = nullptr;
        msplat_hip_gaussians g1 = gauss(nn);
        rc(msplat_hip_render_prior(ctx, &g1, &c, bgv, &hp, &rimg), "render_prior");
        rc(msplat_hip_render_aux_outputs(ctx, &rdep, &rT), "aux outputs");
        Dev fd(nullptr, 4 * (size_t)H * W, ctx), fa(nullptr, 4 * (size_t)H * W, ctx);
        rc(msplat_hip_finalize_depth(ctx, rdep, rT, (uint32_t)(H * W), 0.5f, fd.f(), fa.f()), "finalize");
        std::vector<float> ri(H * W * 3), rfd(H * W), rfa(H * W);
        rc(msplat_hip_memcpy(ctx, ri.data(), rimg, 4 * ri.size(), 1), "image");
        bool pok = std::memcmp(pl, rpl, sizeof pl) == 0;
        for (int k = 0; k < 6; k++) {
            pok = pok && same(bufs[k].data<float>(), dp[k]->get(ctx, (size_t)nn * widths[k]).data(),
                              (size_t)nn * widths[k]);
            pok = pok && same(m2[k].data<float>(), d2[k]->get(ctx, (size_t)nn * widths[k]).data(),
                              (size_t)nn * widths[k]);
        }
        const bool sky_ok = same(sky.data<float>(), dsky.get(ctx, sw * sh * 3).data(), sw * sh * 3) &&
                            same(expo.data<float>(), dexpo.get(ctx, 24).data(), 24) &&
                            same(terms.data<float>(), dterms.get(ctx, 4).data(), 4) &&
                            same(egrad.data<float>(), degrad.get(ctx, 12).data(), 12);
        expect(pok && sky_ok, "prior train_step x2: parameters, moments, sky, exposure, terms");
        expect(same(pimg.data<float>(), ri.data(), ri.size()) &&
                   same(depth_out.data<float>(), fd.get(ctx, H * W).data(), H * W) &&
                   same(alpha_out.data<float>(), fa.get(ctx, H * W).data(), H * W),
               "prior render, aux outputs, finalize_depth");
    }

    for (int k = 0; k < 6; k++) {
        delete dp[k];
        delete d1[k];
        delete d2[k];
    }
    msplat_hip_ctx_destroy(ctx);
    cleanup_msplat_metal();
    std::printf(failures ? "FAILED: %d\n" : "OK\n", failures);
    return failures ? 1 : 0;
}

int main() {
    std::setvbuf(stdout, nullptr, _IOLBF, 0);
    try {
        return run();
    } catch (const std::exception &e) {
        std::printf("exception: %s\n", e.what());
        return 3;
    }
}
