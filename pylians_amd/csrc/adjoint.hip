// Adjoint of particles -> P0(k) and grid interpolation at particle positions, for gfx950.
//
// loss = sum_b w_b P0_b with P0_b = (L/N^2)^3 / Nmodes_b * sum_{independent k in bin b} |c_k delta_k|^2
// (Pk_library.pyx:314-381, 410-417), delta = rho/rhobar - 1.  Then
//   dL/d delta(x) = sum_{all k} H_k e^{i k x},  H_k = w_b (L/N^2)^3 / Nmodes_b c_k^2 delta_k
// over the Hermitian-extended half spectrum, doubled on the self-conjugate modes (counted once in
// the full-space sum but twice in d|delta_k|^2); dL/d rho = dL/d delta / rhobar; and
//   dL/d pos_{p,a} = (N/L) sum_cells G(cell) W'_a prod_{b != a} W_b
// with W the 1-D assignment kernel of the scheme and W' its derivative in cell units.
// CIC_interp (MAS_library.pyx:558-599) is the same gather with W in place of W'.
#include <math.h>

#include "adjoint.hpp"

namespace pyl {

// 1-D kernel and derivative (d/d dist) of the schemes, float64, same stencil cells as deposit.hip
template <int S> struct AxisD;
template <> struct AxisD<1> {
    static __device__ __forceinline__ long long eval(double d, double* w, double* dw) {
        w[0] = 1.0;
        dw[0] = 0.0;
        return (long long)floor(d + 0.5);
    }
};
template <> struct AxisD<2> {
    static __device__ __forceinline__ long long eval(double d, double* w, double* dw) {
        const double b = floor(d);
        const double u = d - b;
        w[0] = 1.0 - u;
        w[1] = u;
        dw[0] = -1.0;
        dw[1] = 1.0;
        return (long long)b;
    }
};
template <> struct AxisD<3> {
    static __device__ __forceinline__ long long eval(double d, double* w, double* dw) {
        const double mn = floor(d - 1.5);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double e = d - (mn + (j + 1));  // signed distance, ds/dd = sign(e)
            const double s = fabs(e), sg = e < 0 ? -1.0 : 1.0;
            if (s < 0.5) {
                w[j] = 0.75 - s * s;
                dw[j] = -2.0 * s * sg;
            } else if (s < 1.5) {
                w[j] = 0.5 * (1.5 - s) * (1.5 - s);
                dw[j] = -(1.5 - s) * sg;
            } else {
                w[j] = 0.0;
                dw[j] = 0.0;
            }
        }
        return (long long)mn + 1;
    }
};
template <> struct AxisD<4> {
    static __device__ __forceinline__ long long eval(double d, double* w, double* dw) {
        const double mn = floor(d - 2.0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double e = d - (mn + (j + 1));
            const double s = fabs(e), sg = e < 0 ? -1.0 : 1.0;
            if (s < 1.0) {
                w[j] = (4.0 - 6.0 * s * s + 3.0 * s * s * s) / 6.0;
                dw[j] = (-12.0 * s + 9.0 * s * s) / 6.0 * sg;
            } else if (s < 2.0) {
                w[j] = (2.0 - s) * (2.0 - s) * (2.0 - s) / 6.0;
                dw[j] = -0.5 * (2.0 - s) * (2.0 - s) * sg;
            } else {
                w[j] = 0.0;
                dw[j] = 0.0;
            }
        }
        return (long long)mn + 1;
    }
};

__device__ __forceinline__ int wrap_i(long long b, int N) {
    long long r = b % N;
    return (int)(r < 0 ? r + N : r);
}

// GRAD: out[3p..3p+2] = scale * dL/dpos;  else out[p] = sum_cells G * W (interpolation)
template <int S, bool GRAD>
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ G, int N, const float* __restrict__ pos,
                                                 long long n, double inv, double scale, float* __restrict__ out) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        double w[3][S], dw[3][S];
        int ix[3][S];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int b = wrap_i(AxisD<S>::eval((double)pos[3 * p + a] * inv, w[a], dw[a]), N);
#pragma unroll
            for (int j = 0; j < S; ++j) ix[a][j] = b + j >= N ? b + j - N : b + j;
        }
        double gx = 0.0, gy = 0.0, gz = 0.0, v = 0.0;
#pragma unroll
        for (int l = 0; l < S; ++l)
#pragma unroll
            for (int m = 0; m < S; ++m) {
                const float* row = G + ((long long)ix[0][l] * N + ix[1][m]) * N;
#pragma unroll
                for (int q = 0; q < S; ++q) {
                    const double g = (double)row[ix[2][q]];
                    if (GRAD) {
                        gx += g * dw[0][l] * w[1][m] * w[2][q];
                        gy += g * w[0][l] * dw[1][m] * w[2][q];
                        gz += g * w[0][l] * w[1][m] * dw[2][q];
                    } else {
                        v += g * w[0][l] * w[1][m] * w[2][q];
                    }
                }
            }
        if (GRAD) {
            out[3 * p] = (float)(gx * scale);
            out[3 * p + 1] = (float)(gy * scale);
            out[3 * p + 2] = (float)(gz * scale);
        } else {
            out[p] = (float)v;
        }
    }
}

template <bool GRAD>
static int launch_gather(const float* G, int N, const float* pos, int64_t n, double inv, double scale, int scheme,
                         float* out, hipStream_t st) {
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    switch (scheme) {
        case 1: hipLaunchKernelGGL((k_gather<1, GRAD>), dim3((unsigned)blocks), dim3(256), 0, st, G, N, pos, (long long)n, inv, scale, out); break;
        case 2: hipLaunchKernelGGL((k_gather<2, GRAD>), dim3((unsigned)blocks), dim3(256), 0, st, G, N, pos, (long long)n, inv, scale, out); break;
        case 3: hipLaunchKernelGGL((k_gather<3, GRAD>), dim3((unsigned)blocks), dim3(256), 0, st, G, N, pos, (long long)n, inv, scale, out); break;
        case 4: hipLaunchKernelGGL((k_gather<4, GRAD>), dim3((unsigned)blocks), dim3(256), 0, st, G, N, pos, (long long)n, inv, scale, out); break;
        default: set_error("unknown scheme %d", scheme); return E_ARG;
    }
    PYL_HIP(hipGetLastError());
    return OK;
}

int gather_grad(const float* G, int N, double BoxSize, const float* pos, int64_t n, int scheme, double scale,
                float* grad, hipStream_t st) {
    if (n <= 0) return OK;
    const double inv = N / BoxSize;
    return launch_gather<true>(G, N, pos, n, inv, scale * inv, scheme, grad, st);
}

int interp(const float* G, int N, double BoxSize, const float* pos, int64_t n, int scheme, float* out,
           hipStream_t st) {
    if (n <= 0) return OK;
    return launch_gather<false>(G, N, pos, n, N / BoxSize, 1.0, scheme, out, st);
}

// H = g(bin) * fac * keep, keep = window-corrected spectrum on independent modes, raw elsewhere
template <typename R>
__global__ __launch_bounds__(256) void k_grad_spec(cplx<R>* __restrict__ spec, int N, long long P,
                                                   const double* __restrict__ w, int nb,
                                                   const unsigned long long* __restrict__ n3, double unit,
                                                   const double* __restrict__ win, double post) {
    const long long total = (long long)N * N * P;
    const int middle = N / 2;
    const bool even = (N % 2) == 0;
    for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < total;
         id += (long long)gridDim.x * blockDim.x) {
        const int kzz = (int)(id % P);
        const long long t = id / P;
        const int yy = (int)(t % N), xx = (int)(t / N);
        if (kzz > middle) continue;
        const int kx = kfreq(xx, N), ky = kfreq(yy, N), kz = kzz;
        const int k2 = kx * kx + ky * ky + kz * kz;   // <= 3 (N/2)^2 < 2^24
        const int bin = isqrt24(k2);
        double g = 0.0;
        if (bin >= 1 && bin <= nb) g = w[bin - 1] * unit / (double)n3[bin];
        const double c = win[xx] * win[yy] * win[kzz];
        bool ind = true;
        if (kz == 0 || (kz == middle && even)) {
            if (kx < 0) ind = false;
            else if ((kx == 0 || (kx == middle && even)) && ky < 0) ind = false;
        }
        double f = g * (ind ? c : c * c) * post;
        if (even && (kx == 0 || kx == middle) && (ky == 0 || ky == middle) && (kz == 0 || kz == middle)) f *= 2.0;
        spec[id] = scale(spec[id], (R)f);
    }
}

template <typename R>
int grad_spectrum(cplx<R>* spec, int N, int64_t P, const double* w, int nb, const unsigned long long* n3,
                  double BoxSize, const double* win, double post, hipStream_t st) {
    const double u1 = BoxSize / ((double)N * N);
    const double unit = u1 * u1 * u1;
    const long long total = (long long)N * N * P;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((k_grad_spec<R>), dim3((unsigned)blocks), dim3(256), 0, st, spec, N, (long long)P, w, nb, n3,
                       unit, win, post);
    PYL_HIP(hipGetLastError());
    return OK;
}
template int grad_spectrum<float>(cplx<float>*, int, int64_t, const double*, int, const unsigned long long*, double,
                                  const double*, double, hipStream_t);
template int grad_spectrum<double>(cplx<double>*, int, int64_t, const double*, int, const unsigned long long*,
                                   double, const double*, double, hipStream_t);

}  // namespace pyl
