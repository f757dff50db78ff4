// Field operations on the same FFT: window correction of a field (Pk_library.correct_MAS), the
// correlation function (Pk_library.Xi) and filter smoothing (smoothing_library FT_filter,
// field_smoothing).
//
// correct_MAS (Pk_library.pyx:1749-1806): r2c, every independent half-spectrum mode multiplied by
//   the float32 window factor, normalised c2r.  The dependent half of the kz = 0 and kz = N/2
//   planes is left as it is, as the reference does.
// Xi (Pk_library.pyx:2035-2151): r2c, every half-spectrum mode replaced by |W delta_k|^2 (float32),
//   normalised c2r, then the real-space field binned by floor(|r|) (cell units) into the
//   multipoles of mu = r_par/|r| with float64 sums.  Cells z and N-z of a row share |r| and mu^2,
//   so they are binned together.
// FT_filter / field_smoothing (smoothing_library.pyx:20-114): the top-hat or Gaussian filter on the
//   grid, normalised by its float64 sum, r2c; field r2c times the filter (complex64), normalised c2r.
#include <math.h>

#include <vector>

#include "../../include/pylians_hip.h"
#include "fft.hpp"
#include "profile.hpp"
#include "util.hpp"

namespace pyl {

const double* pk_window_table(int N, int p);

// ------------------------------------------------------------------ half-spectrum mode kernels
// v *= (float)(W_x W_y W_z); optionally only the independent modes; optionally v = (|v|^2, 0)
template <typename R>
__global__ __launch_bounds__(256) void k_modes(cplx<R>* __restrict__ spec, int N, long long P,
                                               const double* __restrict__ win, int indep_only, int square) {
#pragma clang fp contract(off)
    const int W = N / 2 + 1, middle = N / 2;
    const bool even = (N % 2) == 0;
    const long long total = (long long)N * N * W;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i / ((long long)N * W)), y = (int)((i / W) % N), kz = (int)(i % W);
        if (indep_only && (kz == 0 || (kz == middle && even))) {
            const int kx = kfreq(x, N), ky = kfreq(y, N);
            if (kx < 0 || ((kx == 0 || (kx == middle && even)) && ky < 0)) continue;
        }
        const R f = (R)(win[x] * win[y] * win[kz]);
        cplx<R>* v = spec + ((long long)x * N + y) * P + kz;
        cplx<R> a = *v;
        a.x = a.x * f;
        a.y = a.y * f;
        if (square) {
            const R re = a.x * a.x, im = a.y * a.y;
            a.x = re + im;
            a.y = (R)0;
        }
        *v = a;
    }
}

// spec *= filt (complex multiply in the field precision); both with row pitch P
template <typename R>
__global__ __launch_bounds__(256) void k_cmul(cplx<R>* __restrict__ spec, const cplx<R>* __restrict__ filt, int N,
                                              long long P) {
#pragma clang fp contract(off)
    const int W = N / 2 + 1;
    const long long total = (long long)N * N * W;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const long long row = i / W, kz = i % W;
        const cplx<R> a = spec[row * P + kz], b = filt[row * P + kz];
        const R ac = a.x * b.x, bd = a.y * b.y, ad = a.x * b.y, bc = a.y * b.x;
        spec[row * P + kz] = mk(ac - bd, ad + bc);
    }
}

// ------------------------------------------------------------------ Xi binning
// One wavefront per row (x, y); lane l owns the z pairs u in [l*EU, (l+1)*EU) (z = u and N-u).
// LDS bins (kmax+1) x {sum r, sum xi L0, sum xi L2, sum xi L4, count}; per-block rows of partials.
constexpr int XI_T = 256;
template <typename R>
__global__ __launch_bounds__(XI_T) void k_xi_bin(const R* __restrict__ f, int N, int axis, int nb,
                                                 double* __restrict__ part) {
    extern __shared__ double hx[];   // nb * 5
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < nb * 5; i += XI_T) hx[i] = 0.0;
    __syncthreads();
    const int U = N / 2 + 1;
    const int EU = (U + 63) / 64;
    const long long nrows = (long long)N * N;
    for (long long row = (long long)blockIdx.x * (XI_T / 64) + wave; row < nrows;
         row += (long long)gridDim.x * (XI_T / 64)) {
        const int x = (int)(row / N), y = (int)(row % N);
        const int rx = kfreq(x, N), ry = kfreq(y, N);
        const long long c0 = (long long)rx * rx + (long long)ry * ry;
        const R* fr = f + row * N;
        int cur = -1;
        double sr = 0.0, s0 = 0.0, s2 = 0.0, s4 = 0.0, sc = 0.0;
        for (int q = 0; q < EU; ++q) {
            const int u = lane * EU + q;
            if (u >= U) break;
            const int za = u, zb = (N - u) % N;
            const bool pair = zb != za;
            const double v = (double)fr[za] + (pair ? (double)fr[zb] : 0.0);
            const double m = pair ? 2.0 : 1.0;
            const long long k2 = c0 + (long long)u * u;
            const double k = sqrt((double)k2);
            const int bin = (int)k;
            const double kp = axis == 0 ? (double)rx : (axis == 1 ? (double)ry : (double)u);
            const double mu = k == 0.0 ? 0.0 : kp / k;
            const double mu2 = mu * mu;
            if (bin != cur) {
                if (cur >= 0) {
                    atomicAdd(&hx[5 * cur + 0], sr);
                    atomicAdd(&hx[5 * cur + 1], s0);
                    atomicAdd(&hx[5 * cur + 2], s2);
                    atomicAdd(&hx[5 * cur + 3], s4);
                    atomicAdd(&hx[5 * cur + 4], sc);
                }
                cur = bin;
                sr = s0 = s2 = s4 = sc = 0.0;
            }
            sr += m * k;
            s0 += v;
            s2 += v * (3.0 * mu2 - 1.0) / 2.0;
            s4 += v * (35.0 * mu2 * mu2 - 30.0 * mu2 + 3.0) / 8.0;
            sc += m;
        }
        if (cur >= 0) {
            atomicAdd(&hx[5 * cur + 0], sr);
            atomicAdd(&hx[5 * cur + 1], s0);
            atomicAdd(&hx[5 * cur + 2], s2);
            atomicAdd(&hx[5 * cur + 3], s4);
            atomicAdd(&hx[5 * cur + 4], sc);
        }
    }
    __syncthreads();
    double* rowp = part + (long long)blockIdx.x * nb * 5;
    for (int i = tid; i < nb * 5; i += XI_T) rowp[i] = hx[i];
}

// sum of the per-block rows, one thread per entry (fixed order)
__global__ void k_rows_sum(const double* __restrict__ part, int nblocks, int n, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += part[(long long)b * n + i];
    out[i] = s;
}

// ------------------------------------------------------------------ filters
// Top-hat: 1 where d^2 <= R^2 (cell units); Gaussian: exp(-d^2 / (2 R^2)) (float64, stored float32)
__global__ __launch_bounds__(256) void k_filter(float* __restrict__ out, int N, float R2, int gaussian,
                                                double* __restrict__ part) {
    __shared__ double red[4];
    const int middle = N / 2;
    const long long total = (long long)N * N * N;
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        int a = (int)(i / ((long long)N * N)), b = (int)((i / N) % N), c = (int)(i % N);
        if (a > middle) a -= N;
        if (b > middle) b -= N;
        if (c > middle) c -= N;
        const int d2 = a * a + b * b + c * c;
        float v;
        if (gaussian)
            v = (float)exp(-(double)d2 / (2.0 * (double)R2));
        else
            v = ((float)d2 <= R2) ? 1.0f : 0.0f;
        out[i] = v;
        acc += (double)v;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);   // summed by sum_parts
}

__global__ __launch_bounds__(256) void k_normalise(float* __restrict__ f, long long n, const double* __restrict__ norm) {
    const double s = norm[0];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        f[i] = (float)((double)f[i] / s);
}

// numpy's in-place `x /= d; x -= s` on a grid (Pk_snapshot.py:88, 194, 256): d and s are already
// rounded to the grid type, and each operation rounds once, as numpy's float32 ufuncs do.
template <typename R>
__global__ __launch_bounds__(256) void k_div_sub(R* __restrict__ f, long long n, R d, R s) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const R q = f[i] / d;
        f[i] = q - s;
    }
}

// numpy's `y += a * x` (Pk_snapshot.py:252): the product is rounded before the sum (no fused
// multiply-add), a already rounded to the grid type.
template <typename R>
__global__ __launch_bounds__(256) void k_axpy(R* __restrict__ y, const R* __restrict__ x, long long n, R a) {
#pragma clang fp contract(off)
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const R ax = a * x[i];
        y[i] = y[i] + ax;
    }
}

// ------------------------------------------------------------------ host side
static unsigned grid_for(long long n) {
    long long b = (n + 255) / 256;
    return (unsigned)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

static int check_dims(int dims, int prec) {
    PYL_CHECK(prec == PYL_F32 || prec == PYL_F64, E_ARG, "precision must be PYL_F32 or PYL_F64");
    PYL_CHECK(fft_size_supported(dims), E_UNSUPPORTED, "grid size %d not supported", dims);
    return OK;
}

template <typename R>
static int correct_mas_run(const R* in, R* out, int N, int p, hipStream_t st) {
    int rc = OK;
    const int64_t P = half_len(N);
    const size_t nspec = (size_t)N * N * P;
    cplx<R>* spec = (cplx<R>*)workspace(WS_SPEC, sizeof(cplx<R>) * nspec, &rc);
    if (rc) return rc;
    cplx<R>* ws = (cplx<R>*)workspace(WS_SPEC2, sizeof(cplx<R>) * nspec, &rc);
    if (rc) return rc;
    const double* win = pk_window_table(N, p);
    PYL_CHECK(win, E_ALLOC, "window table allocation failed");
    rc = fft3d_r2c<R>(in, spec, N, P, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_modes<R>, dim3(grid_for((long long)nspec)), dim3(256), 0, st, spec, N, (long long)P, win, 1, 0);
    PYL_HIP(hipGetLastError());
    return fft3d_c2r<R>(spec, out, ws, N, P, (R)(1.0 / ((double)N * N * N)), st);
}

template <typename R>
static int xi_run(const R* in, int N, double BoxSize, int p, int axis, double* out, hipStream_t st) {
    int rc = OK;
    const int64_t P = half_len(N);
    const size_t nspec = (size_t)N * N * P;
    const size_t ncell = (size_t)N * N * N;
    cplx<R>* spec = (cplx<R>*)workspace(WS_SPEC, sizeof(cplx<R>) * nspec, &rc);
    if (rc) return rc;
    cplx<R>* ws = (cplx<R>*)workspace(WS_SPEC2, sizeof(cplx<R>) * nspec, &rc);
    if (rc) return rc;
    R* fxi = (R*)workspace(WS_GRID, sizeof(R) * ncell, &rc);
    if (rc) return rc;
    const double* win = pk_window_table(N, p);
    PYL_CHECK(win, E_ALLOC, "window table allocation failed");
    rc = fft3d_r2c<R>(in, spec, N, P, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_modes<R>, dim3(grid_for((long long)nspec)), dim3(256), 0, st, spec, N, (long long)P, win, 0, 1);
    PYL_HIP(hipGetLastError());
    rc = fft3d_c2r<R>(spec, fxi, ws, N, P, (R)(1.0 / ((double)N * N * N)), st);
    if (rc) return rc;
    const int middle = N / 2;
    int kmax = 0;
    while ((long long)(kmax + 1) * (kmax + 1) <= 3LL * middle * middle) ++kmax;
    const int nb = kmax + 1;
    int nblocks = 2 * cu_count();
    const long long rows = (long long)N * N;
    if ((long long)nblocks * (XI_T / 64) > rows) nblocks = (int)((rows + XI_T / 64 - 1) / (XI_T / 64));
    double* part = (double*)workspace(WS_PART, sizeof(double) * ((size_t)nblocks + 1) * nb * 5, &rc);
    if (rc) return rc;
    double* sums = part + (size_t)nblocks * nb * 5;
    hipLaunchKernelGGL(k_xi_bin<R>, dim3(nblocks), dim3(XI_T), sizeof(double) * nb * 5, st, fxi, N, axis, nb, part);
    PYL_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rows_sum, dim3((nb * 5 + 255) / 256), dim3(256), 0, st, part, nblocks, nb * 5, sums);
    PYL_HIP(hipGetLastError());
    std::vector<double> h((size_t)nb * 5);
    PYL_HIP(hipMemcpyAsync(h.data(), sums, sizeof(double) * nb * 5, hipMemcpyDeviceToHost, st));
    PYL_HIP(hipStreamSynchronize(st));
    // discard the r = 0 bin and give units (Pk_library.pyx:2139-2146)
    const double n3 = (double)N * N * N;
    double* r3D = out;
    double* xi = out + kmax;
    double* nm = out + 4 * (size_t)kmax;
    for (int i = 0; i < kmax; ++i) {
        const double* b = &h[(size_t)5 * (i + 1)];
        const double n = b[4];
        r3D[i] = (b[0] / n) * (BoxSize * 1.0 / N);
        xi[3 * i + 0] = (b[1] / n) * (1.0 / n3);
        xi[3 * i + 1] = (b[2] * 5.0 / n) * (1.0 / n3);
        xi[3 * i + 2] = (b[3] * 9.0 / n) * (1.0 / n3);
        nm[i] = n;
    }
    return OK;
}

}  // namespace pyl

using namespace pyl;

extern "C" int pyl_correct_mas(const void* delta, void* out, int dims, int mas_p, int prec, void* stream) {
    int rc = check_dims(dims, prec);
    if (rc) return rc;
    PYL_CHECK(delta && out, E_ARG, "null pointer");
    PYL_CHECK(mas_p >= 0 && mas_p <= 4, E_ARG, "window exponent must be in [0, 4]");
    hipStream_t st = (hipStream_t)stream;
    StageTimer t("correct_mas", st);
    if (prec == PYL_F32) return correct_mas_run<float>((const float*)delta, (float*)out, dims, mas_p, st);
    return correct_mas_run<double>((const double*)delta, (double*)out, dims, mas_p, st);
}

extern "C" int pyl_xi_layout(int dims, int* kmax, int64_t* out_len) {
    PYL_CHECK(dims >= 2, E_ARG, "bad dims");
    const int middle = dims / 2;
    int k = 0;
    while ((long long)(k + 1) * (k + 1) <= 3LL * middle * middle) ++k;
    if (kmax) *kmax = k;
    if (out_len) *out_len = 5 * (int64_t)k;
    return OK;
}

extern "C" int pyl_xi(const void* delta, int dims, double BoxSize, int mas_p, int axis, int prec, double* out,
                      void* stream) {
    int rc = check_dims(dims, prec);
    if (rc) return rc;
    PYL_CHECK(delta && out, E_ARG, "null pointer");
    PYL_CHECK(mas_p >= 0 && mas_p <= 4, E_ARG, "window exponent must be in [0, 4]");
    PYL_CHECK(axis >= 0 && axis <= 2, E_ARG, "axis must be 0, 1 or 2");
    PYL_CHECK(BoxSize > 0, E_ARG, "BoxSize must be positive");
    hipStream_t st = (hipStream_t)stream;
    StageTimer t("xi", st);
    if (prec == PYL_F32) return xi_run<float>((const float*)delta, dims, BoxSize, mas_p, axis, out, st);
    return xi_run<double>((const double*)delta, dims, BoxSize, mas_p, axis, out, st);
}

extern "C" int pyl_ft_filter(int dims, float BoxSize, float R, int gaussian, void* out_k, void* stream) {
    int rc = check_dims(dims, PYL_F32);
    if (rc) return rc;
    PYL_CHECK(out_k, E_ARG, "null pointer");
    PYL_CHECK(BoxSize > 0.0f, E_ARG, "BoxSize must be positive");
    hipStream_t st = (hipStream_t)stream;
    const size_t ncell = (size_t)dims * dims * dims;
    float* f = (float*)workspace(WS_GRID, sizeof(float) * ncell, &rc);
    if (rc) return rc;
    double* norm = (double*)workspace(WS_MASS, 64, &rc);
    if (rc) return rc;
    const unsigned fb = grid_for((long long)ncell);
    double* part = (double*)workspace(WS_RED, sizeof(double) * fb, &rc);
    if (rc) return rc;
    // R_grid = R*dims/BoxSize and R^2 in float32 (smoothing_library.pyx:34-35)
    const float Rg = (R * (float)dims) / BoxSize;
    const float R2 = Rg * Rg;
    hipLaunchKernelGGL(k_filter, dim3(fb), dim3(256), 0, st, f, dims, R2, gaussian, part);
    PYL_HIP(hipGetLastError());
    rc = sum_parts(part, (int)fb, norm, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_normalise, dim3(grid_for((long long)ncell)), dim3(256), 0, st, f, (long long)ncell, norm);
    PYL_HIP(hipGetLastError());
    return fft3d_r2c<float>(f, (cplx<float>*)out_k, dims, half_len(dims), st);
}

extern "C" int pyl_field_smoothing(const void* field, const void* filter_k, void* out, int dims, void* stream) {
    int rc = check_dims(dims, PYL_F32);
    if (rc) return rc;
    PYL_CHECK(field && filter_k && out, E_ARG, "null pointer");
    hipStream_t st = (hipStream_t)stream;
    const int64_t P = half_len(dims);
    const size_t nspec = (size_t)dims * dims * P;
    cplx<float>* spec = (cplx<float>*)workspace(WS_SPEC, sizeof(cplx<float>) * nspec, &rc);
    if (rc) return rc;
    cplx<float>* ws = (cplx<float>*)workspace(WS_SPEC2, sizeof(cplx<float>) * nspec, &rc);
    if (rc) return rc;
    rc = fft3d_r2c<float>((const float*)field, spec, dims, P, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cmul<float>, dim3(grid_for((long long)nspec)), dim3(256), 0, st, spec,
                       (const cplx<float>*)filter_k, dims, (long long)P);
    PYL_HIP(hipGetLastError());
    return fft3d_c2r<float>(spec, (float*)out, ws, dims, P, (float)(1.0 / ((double)dims * dims * dims)), st);
}

extern "C" int pyl_field_div_sub(void* field, int64_t n, double div, double sub, int prec, void* stream) {
    PYL_CHECK(prec == PYL_F32 || prec == PYL_F64, E_ARG, "precision must be PYL_F32 or PYL_F64");
    PYL_CHECK(n >= 0 && (n == 0 || field), E_ARG, "bad field");
    if (n == 0) return OK;
    hipStream_t st = (hipStream_t)stream;
    if (prec == PYL_F32)
        hipLaunchKernelGGL(k_div_sub<float>, dim3(grid_for(n)), dim3(256), 0, st, (float*)field, (long long)n,
                           (float)div, (float)sub);
    else
        hipLaunchKernelGGL(k_div_sub<double>, dim3(grid_for(n)), dim3(256), 0, st, (double*)field, (long long)n, div,
                           sub);
    PYL_HIP(hipGetLastError());
    return OK;
}

extern "C" int pyl_field_axpy(void* y, const void* x, int64_t n, double a, int prec, void* stream) {
    PYL_CHECK(prec == PYL_F32 || prec == PYL_F64, E_ARG, "precision must be PYL_F32 or PYL_F64");
    PYL_CHECK(n >= 0 && (n == 0 || (x && y)), E_ARG, "bad field");
    if (n == 0) return OK;
    hipStream_t st = (hipStream_t)stream;
    if (prec == PYL_F32)
        hipLaunchKernelGGL(k_axpy<float>, dim3(grid_for(n)), dim3(256), 0, st, (float*)y, (const float*)x,
                           (long long)n, (float)a);
    else
        hipLaunchKernelGGL(k_axpy<double>, dim3(grid_for(n)), dim3(256), 0, st, (double*)y, (const double*)x,
                           (long long)n, a);
    PYL_HIP(hipGetLastError());
    return OK;
}
