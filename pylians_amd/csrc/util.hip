// Error reporting, zero-fill and library-owned workspaces.
#include <stdarg.h>

#include <map>
#include <mutex>

#include "common.hpp"
#include "util.hpp"

namespace pyl {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
const char* last_error() { return g_err; }

// Zero-fill with a kernel rather than hipMemsetAsync, so that the clear is an ordinary kernel
// node when a caller captures the stream into a graph.
__global__ void k_zero(uint4* __restrict__ p, long long n16, unsigned char* __restrict__ tail, int ntail) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
    if (blockIdx.x == 0 && (int)threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

int zero_fill(void* ptr, size_t bytes, hipStream_t st) {
    if (!bytes) return OK;
    PYL_CHECK(((uintptr_t)ptr & 15) == 0, E_ARG, "zero_fill needs 16-byte alignment");
    const long long n16 = (long long)(bytes / 16);
    const int ntail = (int)(bytes % 16);
    long long blocks = (n16 + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, (uint4*)ptr, n16,
                       (unsigned char*)ptr + n16 * 16, ntail);
    PYL_HIP(hipGetLastError());
    return OK;
}

__global__ __launch_bounds__(1024) void k_sum_parts(const double* __restrict__ part, int nb, double* __restrict__ out) {
    __shared__ double red[16];
    const int tid = threadIdx.x;
    double a = 0.0;
    for (int b = tid; b < nb; b += 1024) a += part[b];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = a;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int w = 0; w < 16; ++w) s += red[w];
        out[0] = s;
    }
}

int sum_parts(const double* part, int nb, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1024), 0, st, part, nb, out);
    PYL_HIP(hipGetLastError());
    return OK;
}

// Grow-only device buffers keyed by (device, slot).  Growing synchronises the device before
// freeing the old buffer, so work already queued on it finishes first.
void* workspace(int slot, size_t bytes, int* rc) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::pair<void*, size_t>> ws;
    *rc = OK;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("hipGetDevice failed");
        *rc = E_LAUNCH;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    auto& e = ws[std::make_pair(dev, slot)];
    if (e.second >= bytes && e.first) return e.first;
    if (e.first) {
        hipDeviceSynchronize();
        hipFree(e.first);
        e.first = nullptr;
        e.second = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    if (hipMalloc(&e.first, want) != hipSuccess) {
        e.first = nullptr;
        set_error("workspace allocation of %zu bytes failed", want);
        *rc = E_ALLOC;
        return nullptr;
    }
    e.second = want;
    return e.first;
}

void release_workspaces() {}

}  // namespace pyl
