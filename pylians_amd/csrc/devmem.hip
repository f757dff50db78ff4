// Device memory entry points of the C ABI, for callers that hold their data in host arrays and have
// no device-array library of their own (the numpy drop-in path): allocation, copies, clearing and
// stream synchronisation, each a thin checked wrapper of the HIP runtime call.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/pylians_hip.h"
#include "common.hpp"
#include "util.hpp"

using namespace pyl;

#define PYL_RC(call)                   \
    do {                               \
        const int rc_ = (call);        \
        if (rc_ != OK) return rc_;     \
    } while (0)

namespace {

// Host-to-device copies from pageable memory go through NSTAGE pinned buffers of STAGE bytes owned by
// the library (allocated once, kept until the process ends): host threads copy chunk i into buffer
// i % NSTAGE while the DMA engine moves the chunks before it, so the transfer runs at the pinned rate
// instead of the runtime's pageable path.  The buffer of chunk i is refilled only after the event
// recorded behind its previous DMA has completed.
constexpr size_t STAGE = 32u << 20;
constexpr int NSTAGE = 3;

struct Staging {
    std::mutex mu;
    void* buf[NSTAGE] = {};
    hipEvent_t ev[NSTAGE] = {};
    bool used[NSTAGE] = {};
    bool ready = false;
};

Staging& staging() {
    static Staging s;
    return s;
}

int stage_init(Staging& S) {
    if (S.ready) return OK;
    for (int b = 0; b < NSTAGE; ++b) {
        PYL_HIP(hipHostMalloc(&S.buf[b], STAGE, hipHostMallocDefault));
        PYL_HIP(hipEventCreateWithFlags(&S.ev[b], hipEventDisableTiming));
    }
    S.ready = true;
    return OK;
}

// dst[0, n) = src[0, n) with up to `threads` host threads
void host_copy(char* dst, const char* src, size_t n, int threads) {
    if (threads <= 1 || n < (4u << 20)) {
        memcpy(dst, src, n);
        return;
    }
    // parts of a whole number of pages, threads * part >= n
    const size_t part = ((n + threads - 1) / threads + 4095) & ~(size_t)4095;
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) {
        const size_t a = std::min(n, part * t), b = std::min(n, part * (t + 1));
        if (a < b) pool.emplace_back([=] { memcpy(dst + a, src + a, b - a); });
    }
    memcpy(dst, src, std::min(n, part));
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" int pyl_device_count(int* count) {
    PYL_CHECK(count, E_ARG, "count is NULL");
    *count = 0;
    int n = 0;
    // no device (or no driver) is a count of zero, not an error
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return OK;
    }
    *count = n;
    return OK;
}

extern "C" int pyl_dev_alloc(void** ptr, int64_t bytes) {
    PYL_CHECK(ptr && bytes >= 0, E_ARG, "bad allocation request");
    *ptr = nullptr;
    if (bytes == 0) return OK;
    if (hipMalloc(ptr, (size_t)bytes) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        set_error("device allocation of %lld bytes failed", (long long)bytes);
        return E_ALLOC;
    }
    return OK;
}

extern "C" int pyl_dev_free(void* ptr) {
    if (ptr) PYL_HIP(hipFree(ptr));
    return OK;
}

extern "C" int pyl_memcpy_h2d(void* dst, const void* src, int64_t bytes, void* stream) {
    PYL_CHECK(bytes >= 0 && (bytes == 0 || (dst && src)), E_ARG, "bad copy");
    if (bytes == 0) return OK;
    hipStream_t st = (hipStream_t)stream;
    Staging& S = staging();
    std::lock_guard<std::mutex> lk(S.mu);
    PYL_RC(stage_init(S));
    const unsigned hc = std::thread::hardware_concurrency();
    const int threads = (int)std::max(1u, std::min(8u, hc / 2));
    const char* s = (const char*)src;
    char* d = (char*)dst;
    size_t off = 0;
    for (int i = 0; off < (size_t)bytes; ++i) {
        const int b = i % NSTAGE;
        const size_t n = std::min(STAGE, (size_t)bytes - off);
        if (S.used[b]) PYL_HIP(hipEventSynchronize(S.ev[b]));
        host_copy((char*)S.buf[b], s + off, n, threads);
        PYL_HIP(hipMemcpyAsync(d + off, S.buf[b], n, hipMemcpyHostToDevice, st));
        PYL_HIP(hipEventRecord(S.ev[b], st));
        S.used[b] = true;
        off += n;
    }
    return OK;
}

extern "C" int pyl_memcpy_d2h(void* dst, const void* src, int64_t bytes, void* stream) {
    PYL_CHECK(bytes >= 0 && (bytes == 0 || (dst && src)), E_ARG, "bad copy");
    if (bytes == 0) return OK;
    hipStream_t st = (hipStream_t)stream;
    Staging& S = staging();
    std::lock_guard<std::mutex> lk(S.mu);
    PYL_RC(stage_init(S));
    // host-to-device copies still reading a staging buffer finish before it is written
    for (int b = 0; b < NSTAGE; ++b)
        if (S.used[b]) PYL_HIP(hipEventSynchronize(S.ev[b]));
    const unsigned hc = std::thread::hardware_concurrency();
    const int threads = (int)std::max(1u, std::min(8u, hc / 2));
    const char* s = (const char*)src;
    char* d = (char*)dst;
    const size_t nb = (size_t)bytes, nch = (nb + STAGE - 1) / STAGE;
    // chunk i goes to buffer i % NSTAGE; the DMA of the next NSTAGE - 1 chunks runs while chunk i is
    // copied out by the host threads
    auto issue = [&](size_t i) -> int {
        const int b = (int)(i % NSTAGE);
        const size_t off = i * STAGE, n = std::min(STAGE, nb - off);
        PYL_HIP(hipMemcpyAsync(S.buf[b], s + off, n, hipMemcpyDeviceToHost, st));
        PYL_HIP(hipEventRecord(S.ev[b], st));
        S.used[b] = true;
        return OK;
    };
    for (size_t i = 0; i < std::min(nch, (size_t)NSTAGE); ++i) PYL_RC(issue(i));
    for (size_t i = 0; i < nch; ++i) {
        const int b = (int)(i % NSTAGE);
        const size_t off = i * STAGE, n = std::min(STAGE, nb - off);
        PYL_HIP(hipEventSynchronize(S.ev[b]));
        host_copy(d + off, (const char*)S.buf[b], n, threads);
        if (i + NSTAGE < nch) PYL_RC(issue(i + NSTAGE));
    }
    return OK;
}

extern "C" int pyl_memcpy_d2d(void* dst, const void* src, int64_t bytes, void* stream) {
    PYL_CHECK(bytes >= 0 && (bytes == 0 || (dst && src)), E_ARG, "bad copy");
    if (bytes == 0) return OK;
    PYL_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return OK;
}

extern "C" int pyl_dev_zero(void* ptr, int64_t bytes, void* stream) {
    PYL_CHECK(bytes >= 0 && (bytes == 0 || ptr), E_ARG, "bad clear");
    if (bytes == 0) return OK;
    return zero_fill(ptr, (size_t)bytes, (hipStream_t)stream);
}

extern "C" int pyl_stream_sync(void* stream) {
    PYL_HIP(hipStreamSynchronize((hipStream_t)stream));
    return OK;
}
