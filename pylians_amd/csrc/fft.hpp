// Host-side interface of the 3-D real FFT (r2c / c2r) built from the LDS Stockham engine.
//
// Layout: real grid r[x][y][z] (N^3, z fastest); half spectrum c[x][y][kz], kz in [0, N/2],
// row pitch P (complex elements; P = N/2+1 gives FFTW's / numpy's rfftn layout).
// Forward = unnormalised (FFTW_FORWARD); inverse = unnormalised times `scale`.
#pragma once
#include "common.hpp"

namespace pyl {

bool fft_size_supported(int N);

// compute units of the current device (cached)
int cu_count();

// twiddle table exp(-2 pi i t/N), t in [0, N), cached per (N, precision); device pointer
template <typename R> const cplx<R>* twiddles(int N);

// pass Z: rows r2c. out rows have P complex.  When `affine`, input values are mapped to the
// overdensity v / rhobar - 1 with rhobar = mass / N^3, mass = dmass[0] (device) or `mass`.
template <typename R>
int fft_pass_z_r2c(const R* in, cplx<R>* out, int N, int64_t P, bool affine, const double* dmass, double mass,
                   hipStream_t st);

// column pass along y (axis 1) or x (axis 0) on a half spectrum, src -> dst (may alias).  K > 0:
// transform just the columns kz < K and leave the others untouched.
template <typename R>
int fft_pass_cols(const cplx<R>* src, cplx<R>* dst, int N, int64_t P, int axis, bool inverse, R scale,
                  hipStream_t st, int K = 0);

// pass Z inverse: rows c2r (unnormalised * scale).  store: 0 = nothing written, 1 = x,
// 2 = x * aux.  When acc != nullptr the kernel also adds sum(x^2) into acc[0] and sum(x * aux)
// into acc[1] (float64; aux == nullptr counts as zero).  K > 0: entries kz >= K are zero and
// are not read.
template <typename R>
int fft_pass_z_c2r(const cplx<R>* in, R* out, int N, int64_t P, R scale, int store, const R* aux,
                   double* acc, hipStream_t st, int K = 0);

// Argument blocks of pass Z (rows r2c) and of the column pass, shared with the batch planner
template <typename R> struct R2cArgs {
    const R* in;
    cplx<R>* out;
    long long nrows, P;
    const cplx<R>* tw;
    int affine;
    const double* dmass;
    double mass, ncell;
};

template <typename R> struct ColArgs {
    const cplx<R>* src;
    cplx<R>* dst;
    long long ncols, CW, CS, EST, PW;
    int W;
    const cplx<R>* tw;
    R sc;
};

// Batched launches over jobs of one grid size N (one launch for all of them): *_job builds a job's
// arguments and block count (no launch), *_batch launches the device table tab[0..nj) whose
// block ranges are start[0..nj] (start[nj] = total).
template <typename R>
int r2c_job(const R* in, cplx<R>* out, int N, int64_t P, bool affine, const double* dmass, double mass,
            R2cArgs<R>* a, unsigned* blocks);
template <typename R>
int r2c_batch(int N, const R2cArgs<R>* tab, const unsigned* start, int nj, unsigned total, hipStream_t st);
template <typename R>
int cols_job(const cplx<R>* src, cplx<R>* dst, int N, int64_t P, int axis, ColArgs<R>* a, unsigned* blocks);
template <typename R>
int cols_batch(int N, const ColArgs<R>* tab, const unsigned* start, int nj, unsigned total, hipStream_t st);

// full transforms; with K > 0 only the modes kz < K of the r2c are complete (the column passes
// skip the others, which keep the pass-Z values)
template <typename R>
int fft3d_r2c(const R* in, cplx<R>* out, int N, int64_t P, hipStream_t st, int K = 0);
// in is not modified; ws must hold N*N*P complex values (may equal nothing else)
template <typename R>
int fft3d_c2r(const cplx<R>* in, R* out, cplx<R>* ws, int N, int64_t P, R scale, hipStream_t st);

}  // namespace pyl
