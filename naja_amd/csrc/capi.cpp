// The extern "C" boundary of libnaja_amd.so (declared in include/naja_amd.h).
//
// Argument checks mirror the reference's throws: shape mismatches are
// NAJA_ERR_INVALID_ARGUMENT (backmap.cu:27-35 throws std::invalid_argument),
// an unusable workgroup request is NAJA_ERR_RUNTIME (chr.cu:549 throws
// std::runtime_error("Invalid threads per block")). HIP failures are
// reported as NAJA_ERR_HIP instead of being printed and ignored
// (helper.h:12-27).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/naja_amd.h"
#include "backmap.h"
#include "chr_kernel.h"
#include "common.h"
#include "diag.h"
#include "philox.h"

namespace naja {

namespace {
thread_local char g_error[1024] = {0};
}

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

void clear_error() { g_error[0] = 0; }

namespace {
std::mutex g_pool_mu;
std::vector<hipMemPool_t> g_pools;  // per device, created on first use
}  // namespace

static hipMemPool_t scratch_pool()
{
    int dev = 0;
    NAJA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if ((int)g_pools.size() <= dev) g_pools.resize((size_t)dev + 1, nullptr);
    if (!g_pools[(size_t)dev]) {
        hipMemPoolProps props;
        std::memset(&props, 0, sizeof(props));
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        NAJA_HIP(hipMemPoolCreate(&pool, &props));
        uint64_t keep = kScratchKeepBytes;
        NAJA_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
        g_pools[(size_t)dev] = pool;
    }
    return g_pools[(size_t)dev];
}

void scratch_alloc(void** ptr, size_t bytes, hipStream_t s)
{
    NAJA_HIP(hipMallocFromPoolAsync(ptr, bytes ? bytes : 1, scratch_pool(), s));
}

void scratch_free(void* ptr, hipStream_t s)
{
    if (ptr) NAJA_HIP(hipFreeAsync(ptr, s));
}

HeldBuf& held_buf()
{
    constexpr int kMaxDevices = 64;
    static HeldBuf bufs[kMaxDevices];
    int dev = 0;
    NAJA_HIP(hipGetDevice(&dev));
    NAJA_REQUIRE(dev >= 0 && dev < kMaxDevices, NAJA_ERR_UNSUPPORTED, "device %d", dev);
    return bufs[dev];
}

char* held_reserve(HeldBuf& b, size_t bytes)
{
    if (b.cap < bytes) {
        if (b.dev) NAJA_HIP(hipFree(b.dev));
        b.dev = nullptr;
        b.cap = 0;
        NAJA_HIP(hipMalloc((void**)&b.dev, bytes));
        b.cap = bytes;
    }
    return b.dev;
}

int* held_host(HeldBuf& b, size_t count)
{
    if (b.host_cap < count) {
        if (b.host) NAJA_HIP(hipHostFree(b.host));
        b.host = nullptr;
        b.host_cap = 0;
        NAJA_HIP(hipHostMalloc((void**)&b.host, sizeof(int) * count, hipHostMallocDefault));
        b.host_cap = count;
    }
    return b.host;
}

}  // namespace naja

extern "C" {

const char* naja_last_error(void) { return naja::g_error; }

int naja_version(void) { return 100; }

int naja_backmap(const float* T, const float* shift, const float* Y, float* V, int n, int d, int64_t ns, void* stream)
{
    using namespace naja;
    return guarded([&] {
        NAJA_REQUIRE(T && shift && ((Y && V) || ns == 0), NAJA_ERR_INVALID_ARGUMENT, "backmap_into: NULL operand");
        NAJA_REQUIRE(n > 0 && d > 0 && ns >= 0, NAJA_ERR_INVALID_ARGUMENT,
                     "backmap_into: transformation shape mismatch with samples");
        backmap_launch(T, shift, Y, V, n, d, ns, reinterpret_cast<hipStream_t>(stream));
    });
}

int naja_diag(const float* samples, int dim, int n_total, int nchains, int layout, float* ess, float* rhat,
              float* ess_chain, void* stream)
{
    using namespace naja;
    return guarded([&] {
        NAJA_REQUIRE(samples && ess && rhat, NAJA_ERR_INVALID_ARGUMENT, "NULL samples, ess or rhat");
        NAJA_REQUIRE(dim > 0 && nchains > 0 && n_total > 0, NAJA_ERR_INVALID_ARGUMENT, "invalid diagnostics shape");
        NAJA_REQUIRE(layout == NAJA_LAYOUT_INTERLEAVED || layout == NAJA_LAYOUT_BLOCKS, NAJA_ERR_INVALID_ARGUMENT,
                     "invalid layout %d", layout);
        NAJA_REQUIRE(n_total % nchains == 0, NAJA_ERR_INVALID_ARGUMENT, "n_total (%d) not divisible by n_chains (%d)",
                     n_total, nchains);
        const int spc = n_total / nchains;
        NAJA_REQUIRE(spc >= 4, NAJA_ERR_INVALID_ARGUMENT, "compute_ess: fewer than 4 samples per chain");
        diag_launch(samples, dim, n_total, nchains, layout, ess, rhat, ess_chain, reinterpret_cast<hipStream_t>(stream));
    });
}

int naja_diag_noalloc(const float* samples, int dim, int n_total, int nchains, int layout, float* ess, float* rhat,
                      float* ess_chain, void* stream)
{
    using namespace naja;
    return guarded([&] {
        NAJA_REQUIRE(samples && ess && rhat, NAJA_ERR_INVALID_ARGUMENT, "NULL samples, ess or rhat");
        NAJA_REQUIRE(dim > 0 && nchains > 0 && n_total > 0, NAJA_ERR_INVALID_ARGUMENT, "invalid diagnostics shape");
        NAJA_REQUIRE(layout == NAJA_LAYOUT_INTERLEAVED || layout == NAJA_LAYOUT_BLOCKS, NAJA_ERR_INVALID_ARGUMENT,
                     "invalid layout %d", layout);
        NAJA_REQUIRE(n_total % nchains == 0, NAJA_ERR_INVALID_ARGUMENT, "n_total (%d) not divisible by n_chains (%d)",
                     n_total, nchains);
        NAJA_REQUIRE(n_total / nchains >= 4, NAJA_ERR_INVALID_ARGUMENT, "compute_ess: fewer than 4 samples per chain");
        diag_launch_noalloc(samples, dim, n_total, nchains, layout, ess, rhat, ess_chain,
                            reinterpret_cast<hipStream_t>(stream));
    });
}

int naja_diag_batched(const float* samples, int n_variants, int dim, int n_total, int nchains, int layout, float* ess,
                      float* rhat, float* ess_min, float* rhat_max, void* stream)
{
    using namespace naja;
    return guarded([&] {
        NAJA_REQUIRE(samples && ess && rhat, NAJA_ERR_INVALID_ARGUMENT, "NULL samples, ess or rhat");
        NAJA_REQUIRE(n_variants > 0 && dim > 0 && nchains > 0 && n_total > 0, NAJA_ERR_INVALID_ARGUMENT,
                     "invalid diagnostics shape");
        NAJA_REQUIRE(layout == NAJA_LAYOUT_INTERLEAVED || layout == NAJA_LAYOUT_BLOCKS, NAJA_ERR_INVALID_ARGUMENT,
                     "invalid layout %d", layout);
        NAJA_REQUIRE(n_total % nchains == 0, NAJA_ERR_INVALID_ARGUMENT, "n_total (%d) not divisible by n_chains (%d)",
                     n_total, nchains);
        NAJA_REQUIRE(n_total / nchains >= 4, NAJA_ERR_INVALID_ARGUMENT, "compute_ess: fewer than 4 samples per chain");
        diag_launch(samples, dim, n_total, nchains, layout, ess, rhat, nullptr, reinterpret_cast<hipStream_t>(stream),
                    n_variants, ess_min, rhat_max);
    });
}

}  // extern "C"

extern "C" int naja_chr_timing(int enable)
{
    using namespace naja;
    return guarded([&]() { chr_timing_enable(enable != 0); });
}

extern "C" int naja_chr_kernel_times(int* items, float* ms, int max, int* count)
{
    using namespace naja;
    return guarded([&]() {
        NAJA_REQUIRE(max >= 0 && count, NAJA_ERR_INVALID_ARGUMENT, "naja_chr_kernel_times: bad arguments");
        *count = chr_timing_collect(items, ms, max);
    });
}
