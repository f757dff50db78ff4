// Mass assignment kernels for gfx950.
//
// Stencils (grid points at x_i = i L/N, MAS_library.pyx:123-545, MAS_c.c:8-266):
//   NGP: cell floor(d + 0.5)
//   CIC: cells floor(d), floor(d)+1 with weights 1-u, u (u = d - floor d)
//   TSC: cells floor(d-1.5)+{1,2,3}, w = 0.75-s^2 (s<0.5), 0.5(1.5-s)^2 (s<1.5)
//   PCS: cells floor(d-2)+{1,2,3,4}, w = (4-6s^2+3s^3)/6 (s<1), (2-s)^3/6 (s<2)
// with d = pos * N/L computed in float64 (the reference does this in float32), indices wrapped
// periodically.  The 1-D weights are computed in float64 and rounded once to the grid type.
#include <math.h>

#include "deposit.hpp"
#include "profile.hpp"
#include "util.hpp"

namespace pyl {

template <int S> struct Stencil;

__device__ __forceinline__ int wrap_cell(double c, int dims) {
    long long i = (long long)c % dims;
    return (int)(i < 0 ? i + dims : i);
}

template <> struct Stencil<1> {
    template <typename G> static __device__ __forceinline__ void eval(double d, int dims, int* ix, G* w) {
        ix[0] = wrap_cell(floor(d + 0.5), dims);
        w[0] = (G)1;
    }
};
template <> struct Stencil<2> {
    template <typename G> static __device__ __forceinline__ void eval(double d, int dims, int* ix, G* w) {
        const double b = floor(d);
        const double u = d - b;
        ix[0] = wrap_cell(b, dims);
        ix[1] = ix[0] + 1 == dims ? 0 : ix[0] + 1;
        w[0] = (G)(1.0 - u);
        w[1] = (G)u;
    }
};
template <> struct Stencil<3> {
    template <typename G> static __device__ __forceinline__ void eval(double d, int dims, int* ix, G* w) {
        const double mn = floor(d - 1.5);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double c = mn + (j + 1);
            const double s = fabs(c - d);
            ix[j] = wrap_cell(c, dims);
            w[j] = (G)(s < 0.5 ? 0.75 - s * s : (s < 1.5 ? 0.5 * (1.5 - s) * (1.5 - s) : 0.0));
        }
    }
};
template <> struct Stencil<4> {
    template <typename G> static __device__ __forceinline__ void eval(double d, int dims, int* ix, G* w) {
        const double mn = floor(d - 2.0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double c = mn + (j + 1);
            const double s = fabs(c - d);
            ix[j] = wrap_cell(c, dims);
            w[j] = (G)(s < 1.0 ? (4.0 - 6.0 * s * s + 3.0 * s * s * s) / 6.0
                               : (s < 2.0 ? (2.0 - s) * (2.0 - s) * (2.0 - s) / 6.0 : 0.0));
        }
    }
};

// ------------------------------------------------------------------ direct atomics
template <int S, typename G, bool HASW, int AXES>
__global__ __launch_bounds__(256) void k_deposit_atomic(const float* __restrict__ pos, const float* __restrict__ W,
                                                         long long n, int dims, double inv, G mult,
                                                         G* __restrict__ grid, Rsd rsd) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        int ix[3][S];
        G w[3][S];
#pragma unroll
        for (int a = 0; a < AXES; ++a) {
            float x = pos[p * AXES + a];
            if (rsd.vel && a == rsd.axis) x = rsd_shift(x, rsd.vel[p * 3 + a], rsd.fac, rsd.box);
            Stencil<S>::eval((double)x * inv, dims, ix[a], w[a]);
        }
        const G ww = HASW ? (G)W[p] * mult : mult;
        if (AXES == 3) {
#pragma unroll
            for (int l = 0; l < S; ++l)
#pragma unroll
                for (int m = 0; m < S; ++m) {
                    const G wlm = w[0][l] * w[1][m];
                    const long long row = ((long long)ix[0][l] * dims + ix[1][m]) * dims;
#pragma unroll
                    for (int q = 0; q < S; ++q) atomicAdd(&grid[row + ix[2][q]], wlm * w[2][q] * ww);
                }
        } else {
#pragma unroll
            for (int l = 0; l < S; ++l)
#pragma unroll
                for (int m = 0; m < S; ++m)
                    atomicAdd(&grid[(long long)ix[0][l] * dims + ix[1][m]], w[0][l] * w[1][m] * ww);
        }
    }
}

template <typename G>
__global__ void k_divide(G* __restrict__ g, long long n, G s) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) g[i] /= s;
}

int grid_divide(void* grid, int64_t n, bool f64, double s, hipStream_t st) {
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
    if (f64)
        hipLaunchKernelGGL(k_divide<double>, dim3(blocks), dim3(256), 0, st, (double*)grid, (long long)n, s);
    else
        hipLaunchKernelGGL(k_divide<float>, dim3(blocks), dim3(256), 0, st, (float*)grid, (long long)n, (float)s);
    PYL_HIP(hipGetLastError());
    return OK;
}

// delta = rho * a - 1 with a = ncell / mass, evaluated in float64 and rounded once to float32
// (Pk_snapshot.py:84-88: delta /= mean(delta); delta -= 1).
__global__ void k_delta_round(const double* __restrict__ rho, float* __restrict__ out, long long n,
                              const double* __restrict__ dmass, double mass, double ncell) {
    const double a = ncell / (dmass ? dmass[0] : mass);
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (float)(rho[i] * a - 1.0);
}

// delta = rho * ncell / M - 1 in place on a float64 grid
__global__ void k_delta_inplace(double* __restrict__ g, long long n, const double* __restrict__ dmass, double mass,
                                double ncell) {
    const double a = ncell / (dmass ? dmass[0] : mass);
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) g[i] = g[i] * a - 1.0;
}
int delta_inplace(double* g, int64_t n, const double* dmass, double mass, hipStream_t st) {
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_delta_inplace, dim3(blocks), dim3(256), 0, st, g, (long long)n, dmass, mass, (double)n);
    PYL_HIP(hipGetLastError());
    return OK;
}
int delta_round(const double* rho, float* out, int64_t n, const double* dmass, double mass, hipStream_t st) {
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_delta_round, dim3(blocks), dim3(256), 0, st, rho, out, (long long)n, dmass, mass, (double)n);
    PYL_HIP(hipGetLastError());
    return OK;
}

template <int S, typename G, bool HASW, int AXES>
static int launch_atomic(const float* pos, const float* W, int64_t n, int dims, double inv, double mult, void* grid,
                         hipStream_t st, const Rsd& rsd) {
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL((k_deposit_atomic<S, G, HASW, AXES>), dim3((unsigned)blocks), dim3(256), 0, st, pos, W,
                       (long long)n, dims, inv, (G)mult, (G*)grid, rsd);
    PYL_HIP(hipGetLastError());
    return OK;
}

template <int S, typename G>
static int dispatch_atomic(const float* pos, const float* W, int64_t n, int axes, int dims, double inv, double mult,
                           const Rsd& rsd,
                           void* grid, hipStream_t st) {
    if (axes == 3)
        return W ? launch_atomic<S, G, true, 3>(pos, W, n, dims, inv, mult, grid, st, rsd)
                 : launch_atomic<S, G, false, 3>(pos, W, n, dims, inv, mult, grid, st, rsd);
    return W ? launch_atomic<S, G, true, 2>(pos, W, n, dims, inv, mult, grid, st, Rsd())
             : launch_atomic<S, G, false, 2>(pos, W, n, dims, inv, mult, grid, st, Rsd());
}

template <typename G>
static int dispatch_scheme(const float* pos, const float* W, int64_t n, int axes, int dims, double inv, double mult,
                           const Rsd& rsd,
                           int scheme, void* grid, hipStream_t st) {
    switch (scheme) {
        case 1: return dispatch_atomic<1, G>(pos, W, n, axes, dims, inv, mult, rsd, grid, st);
        case 2: return dispatch_atomic<2, G>(pos, W, n, axes, dims, inv, mult, rsd, grid, st);
        case 3: return dispatch_atomic<3, G>(pos, W, n, axes, dims, inv, mult, rsd, grid, st);
        case 4: return dispatch_atomic<4, G>(pos, W, n, axes, dims, inv, mult, rsd, grid, st);
        default: set_error("unknown mass-assignment scheme %d", scheme); return E_ARG;
    }
}

int deposit(const float* pos, const float* W, int64_t n, int axes, int dims, double BoxSize, int scheme, int flags,
            void* grid, hipStream_t st, const Rsd& rsd) {
    PYL_CHECK(axes == 2 || axes == 3, E_ARG, "axes must be 2 or 3 (got %d)", axes);
    PYL_CHECK(!rsd.vel || (axes == 3 && rsd.axis >= 0 && rsd.axis <= 2), E_ARG, "redshift-space shift needs 3-D positions");
    PYL_CHECK(dims > 0, E_ARG, "dims must be positive");
    PYL_CHECK(BoxSize > 0, E_ARG, "BoxSize must be positive");
    PYL_CHECK(scheme >= 1 && scheme <= 4, E_ARG, "unknown mass-assignment scheme %d", scheme);
    const bool f64 = flags & DEP_GRID_F64;
    int64_t cells = 1;
    for (int a = 0; a < axes; ++a) cells *= dims;
    if (n > 0 && !(flags & DEP_ATOMIC) && binned_supported(axes, dims))
        return deposit_binned(pos, W, n, dims, BoxSize, scheme, f64, (flags & DEP_ZERO) ? 1 : 0, 1.0, grid, nullptr,
                              0.0, st, rsd, (flags & DEP_SORTED) != 0, acc_mode(flags));
    PYL_CHECK(!(flags & (DEP_BINNED | DEP_SORTED)), E_UNSUPPORTED, "binned deposit needs a 3-D grid with dims %% 16 == 0");
    if (flags & DEP_ZERO) {
        StageTimer t("grid_zero", st);
        int rz = zero_fill(grid, (size_t)cells * (f64 ? 8 : 4), st);
        if (rz) return rz;
    }
    const double mult = (axes == 2 && (flags & DEP_DUP2D)) ? (double)scheme : 1.0;
    const double inv = dims / BoxSize;
    int rc = OK;
    if (n > 0) {
        StageTimer t("deposit_atomic", st);
        rc = f64 ? dispatch_scheme<double>(pos, W, n, axes, dims, inv, mult, rsd, scheme, grid, st)
                 : dispatch_scheme<float>(pos, W, n, axes, dims, inv, mult, rsd, scheme, grid, st);
        if (rc) return rc;
    }
    if (axes == 2 && (flags & DEP_DUP2D) && (flags & DEP_RENORM2D) && scheme > 1)
        rc = grid_divide(grid, cells, f64, (double)scheme, st);
    return rc;
}

}  // namespace pyl

namespace pyl {

// float64 sum of the particle weights (the mass that sets rhobar, Pk_snapshot.py:80-84): one
// partial per workgroup, summed by sum_parts
__global__ __launch_bounds__(256) void k_weight_sum(const float* __restrict__ W, long long n, double* __restrict__ part) {
    __shared__ double red[4];
    double a = 0.0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) a += (double)W[i];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

int weight_sum(const float* W, int64_t n, double* out, hipStream_t st) {
    long long blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    int rc = OK;
    double* part = (double*)workspace(WS_RED, sizeof(double) * (size_t)blocks, &rc);
    if (rc) return rc;
    hipLaunchKernelGGL(k_weight_sum, dim3((unsigned)blocks), dim3(256), 0, st, W, (long long)n, part);
    PYL_HIP(hipGetLastError());
    return sum_parts(part, (int)blocks, out, st);
}

// pos_redshift_space in place (redshift_space_library.pyx:29-43): one thread per particle,
// column `axis` of the (n, cols) float32 row-major arrays pos and vel.
__global__ __launch_bounds__(256) void k_pos_rsd(float* __restrict__ pos, const float* __restrict__ vel, long long n,
                                                 int cols, int axis, float fac, float box) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const long long j = i * cols + axis;
        pos[j] = rsd_shift(pos[j], vel[j], fac, box);
    }
}

int pos_redshift_space(float* pos, const float* vel, int64_t n, int cols, int axis, float fac, float box,
                       hipStream_t st) {
    if (n <= 0) return OK;
    long long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_pos_rsd, dim3((unsigned)blocks), dim3(256), 0, st, pos, vel, (long long)n, cols, axis, fac,
                       box);
    PYL_HIP(hipGetLastError());
    return OK;
}

}  // namespace pyl
