// Shared host-side helpers of the naja_amd library: thread-local error state
// and HIP error checks that turn failures into NAJA_ERR_* codes (the
// reference prints CUDA errors and carries on, helper.h:12-27; this library
// reports them to the caller instead).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/naja_amd.h"

namespace naja {

void set_error(const char* fmt, ...);
void clear_error();

struct Error {
    int code;
    explicit Error(int c) : code(c) {}
};

#define NAJA_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t err_ = (call);                                                       \
        if (err_ != hipSuccess) {                                                       \
            ::naja::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(err_),  \
                              __FILE__, __LINE__);                                      \
            throw ::naja::Error(NAJA_ERR_HIP);                                          \
        }                                                                               \
    } while (0)

#define NAJA_REQUIRE(cond, code, ...)          \
    do {                                       \
        if (!(cond)) {                         \
            ::naja::set_error(__VA_ARGS__);    \
            throw ::naja::Error(code);         \
        }                                      \
    } while (0)

// Runs body, mapping naja::Error and std exceptions to return codes.
template <typename F>
int guarded(F&& body)
{
    clear_error();
    try {
        body();
        return NAJA_OK;
    } catch (const Error& e) {
        return e.code;
    } catch (const std::exception& e) {
        set_error("%s", e.what());
        return NAJA_ERR_RUNTIME;
    } catch (...) {
        set_error("unknown C++ exception");
        return NAJA_ERR_RUNTIME;
    }
}

// Stream-ordered device scratch from a library-owned pool (one per device)
// that keeps up to kScratchKeepBytes of freed memory mapped between calls.
// The default pool returns freed memory at every synchronisation, so a call
// that needs a multi-GB buffer would map it again each time (about a
// millisecond per 5 GB). scratch_free(nullptr, s) is a no-op.
constexpr uint64_t kScratchKeepBytes = 16ull << 30;
void scratch_alloc(void** ptr, size_t bytes, hipStream_t s);
void scratch_free(void* ptr, hipStream_t s);
template <typename T>
void scratch_alloc(T** ptr, size_t count, hipStream_t s)
{
    scratch_alloc(reinterpret_cast<void**>(ptr), sizeof(T) * (count ? count : 1), s);
}

// A grow-only device buffer per device, with a pinned host block for status
// read-backs, for calls that synchronise their stream before they return:
// such a call holds the mutex for its whole duration, so no pool call (tens
// to hundreds of microseconds of host time each) is made per call.
struct HeldBuf {
    std::mutex mu;
    char* dev = nullptr;
    size_t cap = 0;
    int* host = nullptr;
    size_t host_cap = 0;
};
HeldBuf& held_buf();                         // the current device's
char* held_reserve(HeldBuf& b, size_t bytes);  // device bytes (old contents dropped)
int* held_host(HeldBuf& b, size_t count);      // pinned host ints

// Runs body(v) for v in [0, n) over up to nthreads host threads.
template <typename F>
void parallel_for(int n, int nthreads, F&& body)
{
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    nthreads = std::max(1, std::min(nthreads, n));
    std::vector<std::thread> pool;
    std::vector<int> codes(nthreads, NAJA_OK);
    std::vector<std::string> msgs(nthreads);
    for (int t = 0; t < nthreads; ++t) {
        pool.emplace_back([&, t]() {
            codes[t] = guarded([&]() {
                for (int v = t; v < n; v += nthreads) body(v);
            });
            if (codes[t] != NAJA_OK) msgs[t] = naja_last_error();
        });
    }
    for (auto& th : pool) th.join();
    for (int t = 0; t < nthreads; ++t)
        if (codes[t] != NAJA_OK) {
            set_error("%s", msgs[t].c_str());
            throw Error(codes[t]);
        }
}

}  // namespace naja
