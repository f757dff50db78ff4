// Shared helpers for the gfx950 kernels of libpylians_hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>
#include <vector>

namespace pyl {

// ---------------------------------------------------------------- errors
// Thread-local last-error string surfaced through pyl_last_error().
void set_error(const char* fmt, ...);
const char* last_error();

#define PYL_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            ::pyl::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,              \
                             hipGetErrorString(e_));                                   \
            return (int)e_ + 1000;                                                     \
        }                                                                              \
    } while (0)

#define PYL_CHECK(cond, code, ...)                                                     \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            ::pyl::set_error(__VA_ARGS__);                                             \
            return (code);                                                             \
        }                                                                              \
    } while (0)

enum Err { OK = 0, E_ARG = 1, E_UNSUPPORTED = 2, E_MODES = 3, E_ALLOC = 4, E_LAUNCH = 5 };

// ---------------------------------------------------------------- complex
template <typename R> struct cplx { R x, y; };

template <typename R> __device__ __forceinline__ cplx<R> mk(R a, R b) { cplx<R> c; c.x = a; c.y = b; return c; }
template <typename R> __device__ __forceinline__ cplx<R> operator+(cplx<R> a, cplx<R> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename R> __device__ __forceinline__ cplx<R> operator-(cplx<R> a, cplx<R> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename R> __device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R> __device__ __forceinline__ cplx<R> scale(cplx<R> a, R s) { return mk(a.x * s, a.y * s); }
template <typename R> __device__ __forceinline__ cplx<R> conjg(cplx<R> a) { return mk(a.x, -a.y); }
// a * (-i)
template <typename R> __device__ __forceinline__ cplx<R> mul_mi(cplx<R> a) { return mk(a.y, -a.x); }
// a * (+i)
template <typename R> __device__ __forceinline__ cplx<R> mul_pi(cplx<R> a) { return mk(-a.y, a.x); }

// exact floor(sqrt(v)) for 0 <= v < 2^52
__device__ __forceinline__ int isqrt_i(long long v) {
    long long r = (long long)sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return (int)r;
}

// floor(sqrt(v)) for 0 <= v < 2^24 (float estimate within one of the root, then corrected)
__device__ __forceinline__ int isqrt24(int v) {
    int r = (int)__builtin_amdgcn_sqrtf((float)v);
    r -= (r * r > v) ? 1 : 0;
    r += ((r + 1) * (r + 1) <= v) ? 1 : 0;
    return r;
}
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }

// signed wavenumber of FFT index i on a grid of n points (Pk_library.pyx:315)
__device__ __forceinline__ int kfreq(int i, int n) { return i > n / 2 ? i - n : i; }

inline int64_t half_len(int n) { return (int64_t)(n / 2 + 1); }

// ---------------------------------------------------------------- block coordinates
// Kernel bodies take their block index and block count explicitly: a single launch passes
// (blockIdx.x, gridDim.x); a batched launch (one launch for the same stage of many jobs) gives every
// job a contiguous range of blocks and passes the index and count within the job.
struct Blk {
    unsigned x, gx;
};
__device__ __forceinline__ Blk own_blk() { return Blk{blockIdx.x, gridDim.x}; }
// job of block b of a batched launch: job j owns blocks [start[j], start[j+1]); start[] ascends and
// every start is a multiple of the XCD count, so a job's block i runs on XCD i % 8 as in a single launch
__device__ __forceinline__ int batch_job(const unsigned* __restrict__ start, int nj, unsigned b) {
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
// the batched counterpart of a body taking (const Args&, Blk): job tab[j], block range from start
#define PYL_BATCH_BLK(start, nj)                                          \
    const int bj_ = ::pyl::batch_job(start, nj, blockIdx.x);              \
    const ::pyl::Blk blk_{blockIdx.x - start[bj_], start[bj_ + 1] - start[bj_]}

}  // namespace pyl
