// Context of the HIP backend (internal): the stream, the grow-only workspace of a
// frame, stage-timing events and the batched-view state.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace msplat {

// Records the message of the last error (msplat_hip_last_error) and returns `code`.
int set_error(int code, const char *fmt, ...);

struct BatchState;
void destroy_batch_state(BatchState *b);

inline int check_camera(const msplat_hip_camera *c) {
    if (!c) return set_error(MSPLAT_ERR_ARG, "camera is NULL");
    if (c->width == 0 || c->height == 0) return set_error(MSPLAT_ERR_ARG, "image size must be positive");
    if (c->tiles_x != (int)((c->width + TILE - 1) / TILE) || c->tiles_y != (int)((c->height + TILE - 1) / TILE))
        return set_error(MSPLAT_ERR_ARG, "tile bounds (%d, %d) do not match %ux%u", c->tiles_x, c->tiles_y, c->width,
                         c->height);
    if (c->tiles_x > 0xFFFF || c->tiles_y > 0xFFFF) return set_error(MSPLAT_ERR_ARG, "image too large");
    if (c->degree > 4) return set_error(MSPLAT_ERR_ARG, "SH degree above 4");
    if (c->degrees_to_use > c->degree)
        return set_error(MSPLAT_ERR_ARG, "degrees_to_use (%u) exceeds the stored SH degree (%u)", c->degrees_to_use,
                         c->degree);
    return MSPLAT_OK;
}

inline int check_gaussians(const msplat_hip_gaussians *g) {
    if (!g) return set_error(MSPLAT_ERR_ARG, "gaussians is NULL");
    if (g->num_points < 0) return set_error(MSPLAT_ERR_ARG, "negative gaussian count");
    if (g->num_points > 0 && (!g->means || !g->scales || !g->quats || !g->features_dc || !g->opacities))
        return set_error(MSPLAT_ERR_ARG, "a gaussian parameter pointer is NULL");
    return MSPLAT_OK;
}

inline int tile_bits(int ntiles) {
    int b = 0;
    while ((1 << b) < ntiles) b++;
    return b;
}


}  // namespace msplat

#define HIPCHECK(expr)                                                                                           \
    do {                                                                                                         \
        hipError_t e_ = (expr);                                                                                  \
        if (e_ != hipSuccess)                                                                                    \
            return msplat::set_error(MSPLAT_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));             \
    } while (0)

#define CHECKED(expr)                     \
    do {                                  \
        int rc_ = (expr);                 \
        if (rc_ != MSPLAT_OK) return rc_; \
    } while (0)

struct Buf {
    void *p = nullptr;
    size_t cap = 0;
};

namespace msplat {
// Replays of the captured one-view train step that may be in flight at once; the inputs
// of each are kept until it is known to have been applied (an over-capacity replay is
// run again from them).
constexpr int STEP_RING = 4;

// The frame state a captured step leaves behind (restored after each replay)
struct StepFrame {
    int n;
    uint32_t width, height;
    int tiles_x, tiles_y, ntiles;
    uint32_t *keys_sorted;
    int32_t *gid_sorted;
    int sort_gid_bits;
    int32_t *vis_sorted;
    int chunks;
    float bg[3];
};

// The inputs of one train step (msplat_hip_train_step's arguments)
struct StepRecord {
    msplat_hip_gaussians g;
    msplat_hip_camera cam;
    float bg[3];
    const float *gt;
    float ssim_weight, inv_n, inv_max;
    msplat_hip_adam adam;
    msplat_hip_grad_stats stats;
    int32_t *radii_out;
};
}  // namespace msplat

struct msplat_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int flags = 0;
    msplat::Tuning tuning;
    // per gaussian
    Buf splat_a, splat_b, splat_blue, depths, aabb, rect, radii, count, offsets, block_sums;
    Buf v_xy, v_conic, v_rgb, v_opac, sh_scratch, big;
    // per intersection
    Buf keys0, keys1, vals0, vals1, slot_a, slot_b, slot_c, slot_f, hist, digit_total, keys64;
    // visible gaussians sorted by depth, per-block counts and scan partials
    Buf block_vis, dk0, dk1, dv0, dv1, ord_sums, scan_part, loff;
    // per tile / pixel
    Buf tile_start, image, final_T, final_idx, v_image, row_loss, loss_scratch;
    Buf grad_scratch;  // stands in for NULL gradient outputs of render_backward
    Buf chunk_prod, chunk_state, chunk_last;  // depth-chunked rasterization planes
    // prior-guided frames: per pixel background, expected-depth sum, dL/dD and dL/dA, the
    // exposure-adjusted render and its gradient; per row prior terms; exposure partials;
    // per intersection depth slots and per gaussian big-record sums of them
    Buf bg_img, out_depth, v_aux, adj, v_adj, prior_rows, expo_part, slot_d, big_d, sky_acc;
    bool frame_aux = false;  // the frame in the workspace is a prior-guided (aux) one
    int frame_chunks = 1;                     // depth chunks per tile of the frame's forward
    Buf aux;           // densify flags/scans/scratch, kNN and metrics workspaces
    int32_t *vis_sorted = nullptr;  // visible gaussians in depth order (the frame's list)
    msplat::BatchState *batch = nullptr;  // batched-view frame and graph (created on first use)
    // captured one-view train step (msplat_hip_set_step_graph)
    bool step_graph_on = false;
    bool capturing = false;         // a capture is being recorded: no allocation, no timing
    uint32_t step_cap_request = 0;  // the caller's intersection capacity (0: from the counts)
    uint64_t step_cap = 0;          // the capacity of the current graph (0: none yet)
    int64_t step_m_hint = 0;        // intersections of the last eager step (the chunk choice)
    hipGraph_t step_graph = nullptr;
    hipGraphExec_t step_exec = nullptr;
    std::vector<uint64_t> step_key;
    Buf step_params;                       // device StepParams the replays read
    msplat::StepParams *h_params = nullptr;  // pinned, STEP_RING slots
    uint32_t *h_flag = nullptr;            // pinned: the overflow word each replay copies back
    hipEvent_t ring_ev[msplat::STEP_RING] = {};
    msplat::StepRecord ring[msplat::STEP_RING];
    uint64_t issued = 0;   // replays since the device's applied count was cleared
    int pending = 0;       // replays not yet settled, the newest `pending` ring entries
    uint64_t next_slot = 0;
    int64_t step_captures = 0, step_replays = 0, step_reruns = 0;
    uint64_t generation = 0;  // bumped whenever a workspace buffer moves
    uint64_t step_gen = 0;    // the generation the current graph was captured at
    msplat::StepFrame step_frame = {};

    // stage timing
    bool timing = false;
    uint32_t stage_mask = ~0u;  // stages timed while timing is on (bit s = stage s)
    struct Mark {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<Mark> marks;
    std::vector<hipEvent_t> event_pool;
    double stage_ms[MSPLAT_NUM_STAGES] = {};
    int64_t stage_calls[MSPLAT_NUM_STAGES] = {};
    int open_stage = -1;
    hipEvent_t open_event = nullptr;

    hipEvent_t take_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    void begin(int stage) {
        if (!timing || capturing || !((stage_mask >> stage) & 1u)) return;
        open_stage = stage;
        open_event = take_event();
        if (open_event) (void)hipEventRecord(open_event, stream);
    }
    void end() {
        if (!timing || capturing || open_stage < 0 || !open_event) return;
        hipEvent_t e = take_event();
        if (!e) return;
        (void)hipEventRecord(e, stream);
        marks.push_back({open_stage, open_event, e});
        open_stage = -1;
        open_event = nullptr;
    }
    void collect() {  // stages whose end event has completed; the others stay open
        std::vector<Mark> open;
        for (auto &m : marks) {
            if (hipEventQuery(m.b) != hipSuccess) {
                open.push_back(m);
                continue;
            }
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
                stage_ms[m.stage] += ms;
                stage_calls[m.stage] += 1;
            }
            event_pool.push_back(m.a);
            event_pool.push_back(m.b);
        }
        marks.swap(open);
    }
    // device scalars: [0] u64 isect total, [8] u32 total, [16] double loss sum, [24] float loss,
    // [32] u32 big count, [40] u64 / [48] u32 visible gaussians, [56] u64 running count of
    // written gradient slots, [64] u32 overflow of a captured step, [68] u32 replays applied
    // since it was cleared, [72] u64 / [80] u64 running counts of the one-view frames'
    // intersections and visible gaussians (zeroed at allocation)
    Buf scalars;
    uint64_t *h_total = nullptr;  // pinned
    hipEvent_t counts_ready = nullptr;  // h_total holds this frame's counts
    // frame state
    int n = 0, ntiles = 0;
    uint32_t width = 0, height = 0;
    int tiles_x = 0, tiles_y = 0;
    int64_t M = 0;
    int64_t V = 0;  // visible gaussians of the frame (tile count > 0)
    bool slot_f_clean = false;  // slot_f[0, M) is all zero: set by the key emission, cleared by a raster backward
    // How far the one-view frame in the workspace is built: the stage entry points refuse
    // to run on a frame whose inputs are missing (they would read unset device pointers).
    enum { FRAME_NONE = 0, FRAME_PROJECTED = 1, FRAME_BINNED = 2, FRAME_RASTERED = 3 };
    int frame_stage = FRAME_NONE;
    uint32_t *keys_sorted = nullptr;  // sorted keys: tile ids, or (tile << sort_gid_bits) | id
    int32_t *gid_sorted = nullptr;    // sorted gaussian ids (null when packed into the keys)
    int sort_gid_bits = 0;            // > 0: the ids are the keys' low sort_gid_bits bits
    float bg[3] = {0, 0, 0};

    int grow(Buf &b, size_t bytes) {
        if (bytes == 0) bytes = 16;
        if (b.cap >= bytes) return MSPLAT_OK;
        if (capturing) return msplat::set_error(MSPLAT_ERR_ARG, "internal: a buffer grows during graph capture");
        generation++;
        if (b.p) {
            HIPCHECK(hipStreamSynchronize(stream));
            HIPCHECK(hipFree(b.p));
            b.p = nullptr;
            b.cap = 0;
        }
        size_t cap = bytes + bytes / 4;  // headroom against regrowth
        HIPCHECK(hipMalloc(&b.p, cap));
        b.cap = cap;
        return MSPLAT_OK;
    }

    void release(Buf &b) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }

    ~msplat_hip_ctx() {
        if (stream) (void)hipStreamSynchronize(stream);
        Buf *all[] = {&splat_a, &splat_b, &splat_blue, &depths, &aabb, &rect, &radii, &count, &offsets, &block_sums,
                      &v_xy, &v_conic, &v_rgb, &v_opac, &sh_scratch, &big, &keys0, &keys1, &vals0, &vals1, &slot_a, &slot_b, &slot_c, &slot_f,
                      &hist, &digit_total, &keys64, &block_vis, &dk0, &dk1, &dv0, &dv1, &ord_sums, &scan_part, &loff, &tile_start, &image, &final_T, &final_idx, &v_image,
                      &row_loss, &loss_scratch, &scalars, &grad_scratch, &aux, &chunk_prod, &chunk_state,
                      &chunk_last, &bg_img, &out_depth, &v_aux, &adj, &v_adj, &prior_rows, &expo_part, &slot_d,
                      &big_d, &sky_acc};
        for (Buf *b : all) release(*b);
        if (batch) msplat::destroy_batch_state(batch);
        if (step_exec) (void)hipGraphExecDestroy(step_exec);
        if (step_graph) (void)hipGraphDestroy(step_graph);
        release(step_params);
        if (h_params) (void)hipHostFree(h_params);
        if (h_flag) (void)hipHostFree(h_flag);
        for (hipEvent_t e : ring_ev)
            if (e) (void)hipEventDestroy(e);
        if (h_total) (void)hipHostFree(h_total);
        if (counts_ready) (void)hipEventDestroy(counts_ready);
        collect();
        for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }

    uint64_t *d_total64() { return (uint64_t *)scalars.p; }
    uint32_t *d_big_count() { return (uint32_t *)((char *)scalars.p + 32); }
    uint32_t *big_list() { return (uint32_t *)big.p; }
    float4 *big_a() { return (float4 *)((char *)big.p + 4 * (size_t)(n > 0 ? n : 1) + 12); }
    uint32_t *d_total32() { return (uint32_t *)((char *)scalars.p + 8); }
    double *d_loss_sum() { return (double *)((char *)scalars.p + 16); }
    uint64_t *d_vis64() { return (uint64_t *)((char *)scalars.p + 40); }
    uint32_t *d_vis32() { return (uint32_t *)((char *)scalars.p + 48); }
    float *d_loss() { return (float *)((char *)scalars.p + 24); }
    unsigned long long *d_slots_written() { return (unsigned long long *)((char *)scalars.p + 56); }
    uint32_t *d_overflow() { return (uint32_t *)((char *)scalars.p + 64); }
    uint32_t *d_applied() { return (uint32_t *)((char *)scalars.p + 68); }
    uint64_t *d_isects_running() { return (uint64_t *)((char *)scalars.p + 72); }
    uint64_t *d_visible_running() { return (uint64_t *)((char *)scalars.p + 80); }
};

namespace msplat {
// Settles the captured train-step replays in flight on the context (waits for them and
// runs again those an intersection overflow left unapplied). Every entry point that reads
// or writes a context's frame, or parameters a step may update, calls it first.
int settle_replays(msplat_hip_ctx *ctx);
}  // namespace msplat
