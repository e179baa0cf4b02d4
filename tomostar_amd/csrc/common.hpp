// Shared host plumbing of libtomostar_hip: the device context, HIP error checks that become status
// codes at the C ABI (nothing throws across it), and an owning device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "tomostar_hip.h"

struct ts_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string name;
    size_t global_mem = 0;
    int cu_count = 0;
    double eik_selftest = NAN;
    double lsqr_selftest = NAN;
    double lsqr_fast_selftest = NAN;
    bool use_graphs = true;  // TOMOSTAR_NO_GRAPHS=1: launch kernels one by one (profilers that trace
                             // graph launches poorly)
    std::mutex gate;  // one solve at a time per handle, as the reference's lock(_gate)
};

namespace ts {

// The stream the calling thread's device work runs on: a serving worker runs a request on its own
// stream whatever context owns the data (a resident matrix, a model), so requests overlap.
inline hipStream_t &stream_override()
{
    static thread_local hipStream_t s = nullptr;
    return s;
}
inline hipStream_t exec_stream(const ts_context *ctx)
{
    hipStream_t s = stream_override();
    return s ? s : ctx->stream;
}

// A failure inside the library; the ABI wrapper turns it into a status and a reason string.
struct error : std::runtime_error {
    int status;
    error(int st, const std::string &m) : std::runtime_error(m), status(st) {}
};

[[noreturn]] inline void fail(int status, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw error(status, buf);
}

#define TS_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            ::ts::fail(TS_EDEVICE, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

inline void set_reason(char *reason, int len, const char *msg)
{
    if (reason && len > 0) {
        std::strncpy(reason, msg, (size_t)len - 1);
        reason[len - 1] = 0;
    }
}

// Runs `body`, mapping library errors and HIP failures to status codes. Device allocation
// failures (out of memory) decline, so the caller falls back, like the reference's buffer errors.
template <class F>
int guarded(char *reason, int len, F &&body)
{
    try {
        set_reason(reason, len, "");
        return body();
    } catch (const error &e) {
        set_reason(reason, len, e.what());
        return e.status;
    } catch (const std::bad_alloc &) {
        set_reason(reason, len, "host memory exhausted");
        return TS_DECLINED;
    } catch (const std::exception &e) {
        set_reason(reason, len, e.what());
        return TS_EDEVICE;
    }
}

// Owning device allocation. Out-of-memory becomes a decline (status 1) with the size in the reason.
template <class T>
struct dbuf {
    T *p = nullptr;
    size_t n = 0;
    dbuf() = default;
    explicit dbuf(size_t count) { alloc(count); }
    dbuf(const dbuf &) = delete;
    dbuf &operator=(const dbuf &) = delete;
    dbuf(dbuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    dbuf &operator=(dbuf &&o) noexcept
    {
        if (this != &o) { reset(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~dbuf() { reset(); }
    void alloc(size_t count)
    {
        reset();
        if (count == 0) return;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            fail(TS_DECLINED, "device buffer of %zu MB: %s", (count * sizeof(T)) >> 20, hipGetErrorString(e));
        }
        n = count;
    }
    void ensure(size_t count)
    {
        if (n < count) alloc(count);
    }
    void reset()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Large device-to-host copies into a caller's (pageable) array: through two pinned staging buffers, the
// device's copy of one chunk overlapping the host threads' copy of the one before. Into pageable memory
// the runtime stages on its own at ~10 GB/s (400 float32 tables of 201 x 201 x 101: 0.63 s, here 0.16 s).
struct HostStage {
    static constexpr size_t kChunk = 128u << 20;
    char *pin[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    HostStage() = default;
    HostStage(const HostStage &) = delete;
    HostStage &operator=(const HostStage &) = delete;
    ~HostStage()
    {
        for (int b = 0; b < 2; b++) {
            if (pin[b]) (void)hipHostFree(pin[b]);
            if (done[b]) (void)hipEventDestroy(done[b]);
        }
    }
};

inline void copy_to_host(std::unique_ptr<HostStage> &hs, hipStream_t st, void *dst, const void *src, size_t bytes)
{
    if (bytes < 2 * HostStage::kChunk) {  // small: the runtime's own copy
        TS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return;
    }
    if (!hs) {
        hs = std::make_unique<HostStage>();
        for (int b = 0; b < 2; b++) {
            TS_HIP(hipHostMalloc((void **)&hs->pin[b], HostStage::kChunk, hipHostMallocDefault));
            TS_HIP(hipEventCreateWithFlags(&hs->done[b], hipEventDisableTiming));
        }
    }
    const size_t chunk = HostStage::kChunk, nchunks = (bytes + chunk - 1) / chunk;
    const unsigned nthreads = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    auto host_copy = [&](size_t c) {  // chunk c, once the device has written it
        const size_t off = c * chunk, len = std::min(chunk, bytes - off), part = (len + nthreads - 1) / nthreads;
        TS_HIP(hipEventSynchronize(hs->done[c % 2]));
        const char *from = hs->pin[c % 2];
        char *to = (char *)dst + off;
        std::vector<std::thread> th;
        for (size_t lo = part; lo < len; lo += part)
            th.emplace_back([=] { std::memcpy(to + lo, from + lo, std::min(part, len - lo)); });
        std::memcpy(to, from, std::min(part, len));
        for (auto &x : th) x.join();
    };
    for (size_t c = 0; c < nchunks; c++) {
        // buffer c % 2 last held chunk c - 2, whose host copy has finished
        const size_t off = c * chunk, len = std::min(chunk, bytes - off);
        TS_HIP(hipMemcpyAsync(hs->pin[c % 2], (const char *)src + off, len, hipMemcpyDeviceToHost, st));
        TS_HIP(hipEventRecord(hs->done[c % 2], st));
        if (c >= 1) host_copy(c - 1);
    }
    host_copy(nchunks - 1);
}

inline bool cancelled(const volatile int32_t *cancel) { return cancel && *cancel != 0; }

}  // namespace ts
