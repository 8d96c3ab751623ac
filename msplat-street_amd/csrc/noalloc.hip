// msplat_hip_radix_sort_noalloc: the stable LSD radix sort of msplat_hip_radix_sort
// without any allocation, for callers that cannot allocate device memory (code under
// stream capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Its scratch (the alternate key and value buffers, the per-block digit histograms, the
// digit totals and the element count) lives in one file-scope device array, `scratch`,
// sized for MSPLAT_NOALLOC_SORT_MAX elements. Every GPU the library runs on holds its
// own copy until the process ends. Host code never passes the array's name to a kernel
// (on the host it is only a stand-in symbol): the first call on a device looks up that
// device's address with hipGetSymbolAddress and keeps it, and every launch takes that
// address. All calls on one GPU share the array, so calls that overlap on one GPU (two
// streams, two threads) race; the kernels fully write the histograms, totals and count
// they read, so nothing needs clearing between calls.
#include <mutex>

#include "context.h"

using namespace msplat;

namespace {

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// scratch layout (bytes): count word | digit totals | histograms | keys | values
constexpr uint32_t kMax = MSPLAT_NOALLOC_SORT_MAX;
constexpr size_t kHistWords = (size_t)256 * ((kMax + 1024 - 1) / 1024 + 1);  // radix_hist_words(kMax)
constexpr size_t kOffTotals = 256;
constexpr size_t kOffHist = align256(kOffTotals + 4 * 256);
constexpr size_t kOffKeys = align256(kOffHist + 4 * kHistWords);
constexpr size_t kOffVals = align256(kOffKeys + 8 * (size_t)kMax);
constexpr size_t kBytes = align256(kOffVals + 4 * (size_t)kMax);

static __device__ __align__(256) unsigned char scratch[kBytes];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on ctx's device, looked up once per device (with
// that device current) and kept.
int scratch_on(const msplat_hip_ctx *ctx, unsigned char **out) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return set_error(MSPLAT_ERR_ARG, "device %d is outside the no-allocation entry point's table", ctx->device);
    std::lock_guard<std::mutex> g(g_lock);
    if (!g_scratch_ptr[ctx->device]) {
        int prev = 0;
        HIPCHECK(hipGetDevice(&prev));
        HIPCHECK(hipSetDevice(ctx->device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        HIPCHECK(hipSetDevice(prev));
        if (e != hipSuccess)
            return set_error(MSPLAT_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        if (!p) return set_error(MSPLAT_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[ctx->device] = p;
    }
    *out = (unsigned char *)g_scratch_ptr[ctx->device];
    return MSPLAT_OK;
}

}  // namespace

static_assert(kHistWords * 4 >= 4 * 256 * ((size_t)(kMax + 1023) / 1024 + 1), "histogram words");

extern "C" {

int msplat_hip_radix_sort_noalloc(msplat_hip_ctx *ctx, uint64_t *keys, int32_t *vals, uint32_t n, int key_bits) {
    if (!ctx) return set_error(MSPLAT_ERR_ARG, "ctx is NULL");
    if (key_bits < 0 || key_bits > 64) return set_error(MSPLAT_ERR_ARG, "key_bits must be in [0, 64]");
    if (n > kMax)
        return set_error(MSPLAT_ERR_RANGE, "radix_sort_noalloc sorts at most %u elements (asked for %u)", kMax, n);
    if (n == 0 || key_bits == 0) return MSPLAT_OK;
    if (!keys || !vals) return set_error(MSPLAT_ERR_ARG, "keys or vals is NULL");
    if (radix_hist_words(n) > kHistWords) return set_error(MSPLAT_ERR_RANGE, "histogram scratch too small");
    unsigned char *base = nullptr;
    CHECKED(scratch_on(ctx, &base));
    uint32_t *dn = (uint32_t *)base;
    uint32_t *totals = (uint32_t *)(base + kOffTotals);
    uint32_t *hist = (uint32_t *)(base + kOffHist);
    fill_u32<<<1, 1, 0, ctx->stream>>>(dn, n);
    uint64_t *kin = keys, *kout = (uint64_t *)(base + kOffKeys);
    int32_t *vin = vals, *vout = (int32_t *)(base + kOffVals);
    for (int shift = 0; shift < key_bits; shift += 8) {
        const int width = std::min(8, key_bits - shift);
        launch_radix_pass(kin, vin, kout, vout, dn, n, shift, hist, totals, ctx->stream, (1u << width) - 1u);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (kin != keys) {
        HIPCHECK(hipMemcpyAsync(keys, kin, 8 * (size_t)n, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHECK(hipMemcpyAsync(vals, vin, 4 * (size_t)n, hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIPCHECK(hipGetLastError());
    return MSPLAT_OK;
}

}  // extern "C"
