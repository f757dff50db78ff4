// Device memory entry points of the C ABI, for callers that hold their data in host arrays and have
// no device-array library of their own (the numpy drop-in path): allocation, copies, clearing and
// stream synchronisation, each a thin checked wrapper of the HIP runtime call.
#include "../../include/pylians_hip.h"
#include "common.hpp"
#include "util.hpp"

using namespace pyl;

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
    PYL_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    return OK;
}

extern "C" int pyl_memcpy_d2h(void* dst, const void* src, int64_t bytes, void* stream) {
    PYL_CHECK(bytes >= 0 && (bytes == 0 || (dst && src)), E_ARG, "bad copy");
    if (bytes == 0) return OK;
    PYL_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    PYL_HIP(hipStreamSynchronize((hipStream_t)stream));
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
