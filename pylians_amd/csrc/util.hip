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
__device__ __forceinline__ void zero_body(uint4* __restrict__ p, long long n16, unsigned char* __restrict__ tail,
                                          int ntail) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
    if (blockIdx.x == 0 && (int)threadIdx.x < ntail) tail[threadIdx.x] = 0;
}
__global__ void k_zero(uint4* __restrict__ p, long long n16, unsigned char* __restrict__ tail, int ntail) {
    zero_body(p, n16, tail, ntail);
}
// batched: buffer blockIdx.y of the table (16-byte aligned pointer, byte count)
__global__ void k_zero_b(const ZeroArgs* __restrict__ tab) {
    const ZeroArgs z = tab[blockIdx.y];
    const long long n16 = (long long)(z.bytes / 16);
    zero_body((uint4*)z.ptr, n16, (unsigned char*)z.ptr + n16 * 16, (int)(z.bytes % 16));
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
struct WsArena {
    int dev = 0;
    bool sealed = false;
    std::map<int, std::pair<void*, size_t>> slots;
    // fixed arena: every slot and table is carved from [base, base + cap) in request order
    char* base = nullptr;
    size_t cap = 0, used = 0;
    hipStream_t st = nullptr;
};
static thread_local WsArena* t_arena = nullptr;

static void* carve(WsArena* a, size_t bytes, int* rc) {
    const size_t at = (a->used + 255) / 256 * 256;
    if (at > a->cap || bytes > a->cap - at) {
        set_error("scratch of %zu bytes exhausted: this call needs more than %zu bytes (a larger grid than the "
                  "no-allocation entry point supports)", a->cap, at + bytes);
        *rc = E_ALLOC;
        return nullptr;
    }
    a->used = at + bytes;
    return a->base + at;
}

static void* grow(std::pair<void*, size_t>& e, size_t bytes, int* rc) {
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
    if (WsArena* a = t_arena) {
        if (a->dev != dev) {
            set_error("graph workspace of device %d used on device %d", a->dev, dev);
            *rc = E_ARG;
            return nullptr;
        }
        auto& e = a->slots[slot];
        if (e.second >= bytes && e.first) return e.first;
        if (a->base) {
            void* p = carve(a, bytes, rc);
            if (p) e = std::make_pair(p, bytes);
            return p;
        }
        if (a->sealed) {
            set_error("graph workspace slot %d would grow from %zu to %zu bytes during capture", slot, e.second,
                      bytes);
            *rc = E_ALLOC;
            return nullptr;
        }
        return grow(e, bytes, rc);
    }
    std::lock_guard<std::mutex> lk(mu);
    auto& e = ws[std::make_pair(dev, slot)];
    if (e.second >= bytes && e.first) return e.first;
    return grow(e, bytes, rc);
}

void* host_pinned(size_t bytes, int* rc) {
    static std::mutex mu;
    static std::map<int, std::pair<void*, size_t>> bufs;
    *rc = OK;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("hipGetDevice failed");
        *rc = E_LAUNCH;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    auto& e = bufs[dev];
    if (e.first && e.second >= bytes) return e.first;
    if (e.first) {
        (void)hipDeviceSynchronize();
        (void)hipHostFree(e.first);
        e = std::make_pair(nullptr, (size_t)0);
    }
    const size_t want = bytes < 4096 ? 4096 : bytes;
    if (hipHostMalloc(&e.first, want, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        e.first = nullptr;
        set_error("pinned host allocation of %zu bytes failed", want);
        *rc = E_ALLOC;
        return nullptr;
    }
    e.second = want;
    return e.first;
}

WsArena* ws_arena_create() {
    WsArena* a = new WsArena();
    if (hipGetDevice(&a->dev) != hipSuccess) a->dev = 0;
    return a;
}

WsArena* ws_arena_fixed(void* base, size_t cap) {
    WsArena* a = ws_arena_create();
    a->base = (char*)base;
    a->cap = cap;
    return a;
}

void ws_arena_fixed_reset(WsArena* a, hipStream_t st) {
    a->slots.clear();
    a->used = 0;
    a->st = st;
}

bool arena_fixed() { return t_arena && t_arena->base; }

void* arena_upload(const void* host, size_t bytes, int* rc) {
    *rc = OK;
    WsArena* a = t_arena;
    if (!a || !a->base) {
        set_error("arena_upload outside a fixed arena");
        *rc = E_ARG;
        return nullptr;
    }
    void* p = carve(a, bytes, rc);
    if (!p) return nullptr;
    if (hipMemcpyAsync(p, host, bytes, hipMemcpyHostToDevice, a->st) != hipSuccess) {
        set_error("table upload into the scratch failed");
        *rc = E_LAUNCH;
        return nullptr;
    }
    return p;
}

void ws_arena_seal(WsArena* a) {
    if (a) a->sealed = true;
}

void ws_arena_destroy(WsArena* a) {
    if (!a) return;
    if (a->base) {   // fixed arenas own no memory
        delete a;
        return;
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(a->dev);
    (void)hipDeviceSynchronize();
    for (auto& kv : a->slots)
        if (kv.second.first) (void)hipFree(kv.second.first);
    (void)hipSetDevice(cur);
    delete a;
}

size_t ws_arena_bytes(const WsArena* a) {
    size_t s = 0;
    if (a)
        for (auto& kv : a->slots) s += kv.second.second;
    return s;
}

ArenaScope::ArenaScope(WsArena* a) : prev(t_arena) { t_arena = a; }
ArenaScope::~ArenaScope() { t_arena = prev; }

}  // namespace pyl
