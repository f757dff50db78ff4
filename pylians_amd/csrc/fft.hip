// 3-D real FFT passes for gfx950: rows r2c/c2r (pass Z) and column passes (Y, X).
#include <math.h>

#include <map>
#include <mutex>
#include <vector>

#include "fft.hpp"
#include "fft_engine.hpp"
#include "util.hpp"

namespace pyl {

// ----------------------------------------------------------------------- kernels
// Pass Z forward: NL rows of N = 2M reals, packed as M complex z[j] = x[2j] + i x[2j+1],
// M-point FFT in LDS, then X[k] = E[k] + W_N^k O[k] for k in [0, M].  Persistent blocks walk row
// blocks; the next block's rows are loaded (coalesced along the rows) into registers while the
// current one is transformed.

template <typename R, int M, int NL, int T>
__device__ __forceinline__ void r2c_rows_body(const R2cArgs<R>& A, Blk B) {
    const R* __restrict__ in = A.in;
    cplx<R>* __restrict__ out = A.out;
    const long long nrows = A.nrows, P = A.P;
    const cplx<R>* __restrict__ twg = A.tw;
    const int affine = A.affine;
    const double* __restrict__ dmass = A.dmass;
    const double mass = A.mass, ncell = A.ncell;
    constexpr int N = 2 * M;
    constexpr int ES = NL + 1;
    constexpr int V = (NL * M + T - 1) / T;   // complex inputs per thread per row block
    __shared__ cplx<R> s[M * ES];
    __shared__ cplx<R> tw[N];
    const int tid = threadIdx.x;
    // overdensity delta = rho/rhobar - 1, rhobar = mass/ncell (Pk_snapshot.py:84-88)
    const R a = affine ? (R)(ncell / (dmass ? dmass[0] : mass)) : (R)1;
    const R b = (R)-1;
    const long long nrb = (nrows + NL - 1) / NL;
    for (int i = tid; i < N; i += T) tw[i] = twg[i];
    cplx<R> nxt[V];
    auto load = [&](long long rb) {
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const int idx = tid + q * T;
            const int r = idx / M, j = idx % M;
            const long long row = rb * NL + r;
            cplx<R> v = mk<R>(0, 0);
            if ((NL * M) % T == 0 || idx < NL * M) {
                if (row < nrows) {
                    const R* p = in + row * N + 2 * j;
                    v = mk(p[0], p[1]);
                }
            }
            nxt[q] = v;
        }
    };
    long long rb = B.x;
    if (rb < nrb) load(rb);
    for (; rb < nrb; rb += B.gx) {
        cplx<R> cur[V];
#pragma unroll
        for (int q = 0; q < V; ++q) cur[q] = nxt[q];
        if (rb + B.gx < nrb) load(rb + B.gx);
        __syncthreads();   // the previous row block's post-processing has read s
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const int idx = tid + q * T;
            if ((NL * M) % T == 0 || idx < NL * M) {
                const int r = idx / M, j = idx % M;
                cplx<R> v = cur[q];
                if (affine) {
                    v.x = v.x * a + b;
                    v.y = v.y * a + b;
                }
                s[j * ES + r] = v;
            }
        }
        __syncthreads();
        lds_fft<R, M, NL, ES, T, 2>(s, tw, tid);
        const long long row0 = rb * NL;
        for (int idx = tid; idx < NL * (M + 1); idx += T) {
            const int r = idx / (M + 1), k = idx % (M + 1);
            const long long row = row0 + r;
            if (row >= nrows) continue;
            const cplx<R> zk = s[(k % M) * ES + r];
            const cplx<R> zm = conjg(s[((M - k) % M) * ES + r]);
            const cplx<R> e = scale(zk + zm, (R)0.5);
            const cplx<R> o = scale(mul_mi(zk - zm), (R)0.5);
            out[row * P + k] = e + cmul(tw[k], o);
        }
    }
}
template <typename R, int M, int NL, int T>
__global__ __launch_bounds__(T) void k_r2c_rows(R2cArgs<R> a) {
    r2c_rows_body<R, M, NL, T>(a, own_blk());
}
template <typename R, int M, int NL, int T>
__global__ __launch_bounds__(T) void k_r2c_rows_b(const R2cArgs<R>* __restrict__ tab, const unsigned* __restrict__ start,
                                                   int nj) {
    PYL_BATCH_BLK(start, nj);
    r2c_rows_body<R, M, NL, T>(tab[bj_], blk_);
}

// Pass Z inverse: rows of M+1 complex -> N = 2M reals.
// Z'[k] = (A + B) + i e^{+2 pi i k/N} (A - B), A = X[k], B = conj(X[M-k]); z = IFFT_M(Z');
// x[2j] = Re z[j], x[2j+1] = Im z[j].  The inverse is computed as conj(FFT(conj(Z'))).
// store: 0 = no output, 1 = x, 2 = x * aux.  Optional fused float64 reductions: each workgroup
// writes its sum x^2 and sum x*aux to part[2 * block + 0 / 1] (k_c2r_partsum adds them up).
template <typename R, int M, int NL, int T>
__global__ __launch_bounds__(T) void k_c2r_rows(const cplx<R>* __restrict__ in, R* __restrict__ out,
                                                 long long nrows, long long P,
                                                 const cplx<R>* __restrict__ twg, R sc, int store,
                                                 const R* __restrict__ aux, double* __restrict__ part, int K) {
    constexpr int N = 2 * M;
    constexpr int ES = NL + 1;
    // whole wavefronts only: the float64 sums below reduce across all 64 lanes of every wave
    static_assert(T % 64 == 0, "k_c2r_rows needs a multiple of 64 threads");
    __shared__ cplx<R> s[M * ES];
    __shared__ cplx<R> tw[N];
    __shared__ double red[2][T / 64];
    const int tid = threadIdx.x;
    const long long row0 = (long long)blockIdx.x * NL;
    for (int i = tid; i < N; i += T) tw[i] = twg[i];
    __syncthreads();
    for (int idx = tid; idx < NL * M; idx += T) {
        const int r = idx / M, k = idx % M;
        const long long row = row0 + r;
        cplx<R> v = mk<R>(0, 0);
        if (row < nrows) {
            // entries kz >= K are known zeros (K = M + 1: the whole row) and are not read
            cplx<R> A = k < K ? in[row * P + k] : mk<R>(0, 0);
            cplx<R> B = M - k < K ? conjg(in[row * P + (M - k)]) : mk<R>(0, 0);
            // the imaginary parts of the DC and Nyquist terms of a row do not exist in a real
            // transform: drop them, as numpy's irfftn and FFTW's c2r do for non-Hermitian input
            if (k == 0) {
                A.y = (R)0;
                B.y = (R)0;
            }
            const cplx<R> e = A + B;
            const cplx<R> o = cmul(A - B, conjg(tw[k]));
            v = conjg(e + mul_pi(o));
        }
        s[k * ES + r] = v;
    }
    __syncthreads();
    lds_fft<R, M, NL, ES, T, 2>(s, tw, tid);
    double a0 = 0.0, a1 = 0.0;
    for (int idx = tid; idx < NL * M; idx += T) {
        const int r = idx / M, j = idx % M;
        const long long row = row0 + r;
        if (row >= nrows) continue;
        const cplx<R> z = s[j * ES + r];
        const R x0 = z.x * sc, x1 = -z.y * sc;
        R q0 = (R)0, q1 = (R)0;
        if (aux) {
            const R* q = aux + row * N + 2 * j;
            q0 = q[0];
            q1 = q[1];
        }
        if (store) {
            R* p = out + row * N + 2 * j;
            p[0] = store == 2 ? x0 * q0 : x0;
            p[1] = store == 2 ? x1 * q1 : x1;
        }
        if (part) {
            a0 += (double)x0 * x0 + (double)x1 * x1;
            a1 += (double)x0 * q0 + (double)x1 * q1;
        }
    }
    if (part) {
        for (int off = 32; off > 0; off >>= 1) {
            a0 += __shfl_down(a0, off, 64);
            a1 += __shfl_down(a1, off, 64);
        }
        if ((tid & 63) == 0) {
            red[0][tid >> 6] = a0;
            red[1][tid >> 6] = a1;
        }
        __syncthreads();
        if (tid == 0) {
            double s0 = 0.0, s1 = 0.0;
            for (int w = 0; w < T / 64; ++w) {
                s0 += red[0][w];
                s1 += red[1][w];
            }
            part[2 * blockIdx.x] = s0;
            part[2 * blockIdx.x + 1] = s1;
        }
    }
}

// acc[0] += sum of part[2 b], acc[1] += sum of part[2 b + 1] (when two) over the nb workgroups of
// k_c2r_rows; one workgroup, one atomic per sum
__global__ __launch_bounds__(1024) void k_c2r_partsum(const double* __restrict__ part, int nb, double* __restrict__ acc,
                                                      int two) {
    __shared__ double red[2][16];
    const int tid = threadIdx.x;
    double a0 = 0.0, a1 = 0.0;
    for (int b = tid; b < nb; b += 1024) {
        a0 += part[2 * b];
        a1 += part[2 * b + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_down(a0, off, 64);
        a1 += __shfl_down(a1, off, 64);
    }
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = a0;
        red[1][tid >> 6] = a1;
    }
    __syncthreads();
    if (tid == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int w = 0; w < 16; ++w) {
            s0 += red[0][w];
            s1 += red[1][w];
        }
        atomicAdd(acc, s0);
        if (two) atomicAdd(acc + 1, s1);
    }
}

// Column FFT over a half spectrum with row pitch P (kz < W = N/2+1 hold data).  Column c (of
// ncols) starts at (c / CW) * CS + c % CW and its N elements are EST apart.  Pass Y: CW = P,
// CS = N*P, EST = P.  Pass X: CW = ncols, EST = N*P.  Columns with c % P >= W are skipped.
// Persistent blocks walk column blocks of NL columns; while one block is transformed the first-stage
// inputs of the next one are already in flight into registers, and the last stage stores straight
// from registers.  With a single radix stage the transform never touches LDS.

template <typename R, int N, int NL, int T, bool INV>
__device__ __forceinline__ void cols_body(const ColArgs<R>& A, Blk B) {
    const cplx<R>* src = A.src;
    cplx<R>* dst = A.dst;
    const long long ncols = A.ncols, CW = A.CW, CS = A.CS, EST = A.EST, PW = A.PW;
    const int W = A.W;
    const cplx<R>* __restrict__ twg = A.tw;
    const R sc = A.sc;
    static_assert(T % NL == 0, "T must be a multiple of NL");
    using F = First<N, NL, T>;
    // prefetching doubles the first stage's registers; beyond 512 points it would spill
    constexpr bool PREFETCH = N <= 512;
    __shared__ cplx<R> s[N * NL];
    __shared__ cplx<R> tw[N];
    const int tid = threadIdx.x;
    for (int i = tid; i < N; i += T) tw[i] = twg[i];
    const int l = tid % NL;   // every butterfly of this thread belongs to column l of the block
    const long long nblk = (ncols + NL - 1) / NL;
    auto col_base = [&](long long cb, bool& valid) {
        const long long c = cb * NL + l;
        valid = c < ncols && (int)(c % PW) < W;   // kz = c % PW; columns kz >= W are row padding
        return valid ? (c / CW) * CS + (c % CW) : 0;
    };
    cplx<R> nxt[F::B][F::RAD];
    long long cb = B.x;
    if (PREFETCH && cb < nblk) {
        bool v;
        const long long b = col_base(cb, v);
        first_load<R, N, NL, T>(nxt, [&](int j, int) {
            const cplx<R> x = v ? src[b + j * EST] : mk<R>(0, 0);
            return INV ? conjg(x) : x;
        }, tid);
    }
    for (; cb < nblk; cb += B.gx) {
        cplx<R> cur[F::B][F::RAD];
        if (PREFETCH) {
#pragma unroll
            for (int b = 0; b < F::B; ++b)
#pragma unroll
                for (int m = 0; m < F::RAD; ++m) cur[b][m] = nxt[b][m];
        } else {
            bool v;
            const long long b = col_base(cb, v);
            first_load<R, N, NL, T>(cur, [&](int j, int) {
                const cplx<R> x = v ? src[b + j * EST] : mk<R>(0, 0);
                return INV ? conjg(x) : x;
            }, tid);
        }
        const long long cbn = cb + B.gx;
        if (PREFETCH && cbn < nblk) {
            bool v;
            const long long b = col_base(cbn, v);
            first_load<R, N, NL, T>(nxt, [&](int j, int) {
                const cplx<R> x = v ? src[b + j * EST] : mk<R>(0, 0);
                return INV ? conjg(x) : x;
            }, tid);
        }
        bool valid;
        const long long base = col_base(cb, valid);
        auto store = [&](int k, int, cplx<R> x) {
            if (valid) dst[base + k * EST] = scale(INV ? conjg(x) : x, sc);
        };
        if constexpr (F::NST == 1) {
#pragma unroll
            for (int b = 0; b < F::B; ++b) {
                const int g = tid + b * T;
                if (F::NBF % T == 0 || g < F::NBF) {
                    const int j = g / NL;
                    Dft<R, F::RAD>::run(cur[b]);
#pragma unroll
                    for (int m = 0; m < F::RAD; ++m) store(j * F::RAD + m, l, cur[b][m]);
                }
            }
        } else {
            __syncthreads();  // LDS of the previous block consumed (and twiddles loaded)
            first_stage_regs<R, N, NL, NL, T>(s, cur, tid);
            mid_stages<R, N, NL, NL, T, 1, 1, First<N, NL, T>::RAD>(s, tw, tid);
            last_stage_store<R, N, NL, NL, T, 1>(s, tw, store, tid);
        }
    }
}
template <typename R, int N, int NL, int T, bool INV>
__global__ __launch_bounds__(T) void k_cols(ColArgs<R> a) {
    cols_body<R, N, NL, T, INV>(a, own_blk());
}
template <typename R, int N, int NL, int T, bool INV>
__global__ __launch_bounds__(T) void k_cols_b(const ColArgs<R>* __restrict__ tab, const unsigned* __restrict__ start,
                                               int nj) {
    PYL_BATCH_BLK(start, nj);
    cols_body<R, N, NL, T, INV>(tab[bj_], blk_);
}

// ----------------------------------------------------------------------- dispatch
#define PYL_FFT_SIZES(X) \
    X(8) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160) X(192) X(256) \
    X(320) X(384) X(512) X(640) X(768) X(1024)

bool fft_size_supported(int N) {
    switch (N) {
#define PYL_CASE(n) case n:
        PYL_FFT_SIZES(PYL_CASE)
#undef PYL_CASE
        return true;
        default:
            return false;
    }
}

int cu_count() {
    static int n = 0;
    if (!n) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            n = v;
        else
            n = 256;
    }
    return n;
}

// host table exp(-2 pi i t / N), t < N, rounded once from long double; cached per N
template <typename R> static const std::vector<cplx<R>>& host_twiddles(int N) {
    static std::mutex mu;
    static std::map<int, std::vector<cplx<R>>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(N);
    if (it != cache.end()) return it->second;
    std::vector<cplx<R>>& h = cache[N];
    h.resize(N);
    for (int t = 0; t < N; ++t) {
        long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)N;
        h[t].x = (R)cosl(ang);
        h[t].y = (R)sinl(ang);
    }
    return h;
}

// device copy: cached per (device, N); inside a fixed arena a fresh copy in the arena
template <typename R> const cplx<R>* twiddles(int N) {
    const std::vector<cplx<R>>& h = host_twiddles<R>(N);
    if (arena_fixed()) {
        int rc = OK;
        return (const cplx<R>*)arena_upload(h.data(), sizeof(cplx<R>) * N, &rc);
    }
    static std::mutex mu;
    static std::map<std::pair<int, int>, void*> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(dev, N);
    auto it = cache.find(key);
    if (it != cache.end()) return (const cplx<R>*)it->second;
    void* d = nullptr;
    if (hipMalloc(&d, sizeof(cplx<R>) * N) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), sizeof(cplx<R>) * N, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    cache[key] = d;
    return (const cplx<R>*)d;
}
template const cplx<float>* twiddles<float>(int);
template const cplx<double>* twiddles<double>(int);

// pass Z launch shape of one job: M = N/2-point complex rows, NL rows per block; persistent grid of
// about as many blocks as fit on the chip at once (LDS bound)
template <typename R, int N> struct R2cShape {
    static constexpr int M = N / 2;
    static constexpr int NL = Geo<R>::NL;
    // rows with a radix-3 factor: 8 instead of 16 complex values per thread measured faster
    static constexpr int T = threads_for(M, NL) * ((M % 3 == 0 && M >= 96) ? 2 : 1);
    static constexpr int LDS = (M * (NL + 1) + N) * (int)sizeof(cplx<R>);
};

template <typename R, int N>
static int r2c_job_n(const R* in, cplx<R>* out, long long nrows, int64_t P, bool affine, const double* dmass,
                     double mass, R2cArgs<R>* a, unsigned* blocks) {
    using SH = R2cShape<R, N>;
    const cplx<R>* tw = twiddles<R>(N);
    PYL_CHECK(tw, E_ALLOC, "twiddle table allocation failed");
    int per_cu = (160 * 1024) / (SH::LDS + 1024);
    per_cu = per_cu < 1 ? 1 : (per_cu > 8 ? 8 : per_cu);
    long long nb = (long long)cu_count() * per_cu;
    const long long nrb = (nrows + SH::NL - 1) / SH::NL;
    if (nb > nrb) nb = nrb;
    *a = R2cArgs<R>{in, out, nrows, (long long)P, tw, affine ? 1 : 0, dmass, mass, (double)N * N * N};
    *blocks = (unsigned)nb;
    return OK;
}

template <typename R, int N>
static int launch_r2c_rows(const R* in, cplx<R>* out, long long nrows, int64_t P, bool affine,
                           const double* dmass, double mass, hipStream_t st) {
    using SH = R2cShape<R, N>;
    R2cArgs<R> a;
    unsigned nb = 0;
    int rc = r2c_job_n<R, N>(in, out, nrows, P, affine, dmass, mass, &a, &nb);
    if (rc) return rc;
    hipLaunchKernelGGL((k_r2c_rows<R, SH::M, SH::NL, SH::T>), dim3(nb), dim3(SH::T), 0, st, a);
    PYL_HIP(hipGetLastError());
    return OK;
}

template <typename R, int N>
static int launch_r2c_batch(const R2cArgs<R>* tab, const unsigned* start, int nj, unsigned total, hipStream_t st) {
    using SH = R2cShape<R, N>;
    hipLaunchKernelGGL((k_r2c_rows_b<R, SH::M, SH::NL, SH::T>), dim3(total), dim3(SH::T), 0, st, tab, start, nj);
    PYL_HIP(hipGetLastError());
    return OK;
}

template <typename R, int N>
static int launch_c2r_rows(const cplx<R>* in, R* out, long long nrows, int64_t P, R sc, int store,
                           const R* aux, double* acc, hipStream_t st, int K) {
    constexpr int M = N / 2;
    constexpr int NL = Geo<R>::NL;
    // as pass Z forward, rounded up to whole wavefronts (the fused sums reduce over full waves):
    // M = 80 or 160 points would otherwise give 80 or 160 threads
    constexpr int T = (threads_for(M, NL) * ((M % 3 == 0 && M >= 96) ? 2 : 1) + 63) / 64 * 64;
    const cplx<R>* tw = twiddles<R>(N);
    PYL_CHECK(tw, E_ALLOC, "twiddle table allocation failed");
    const long long nb = (nrows + NL - 1) / NL;
    double* part = nullptr;
    if (acc) {
        int rc = OK;
        part = (double*)workspace(WS_C2R, sizeof(double) * 2 * (size_t)nb, &rc);
        if (rc) return rc;
    }
    hipLaunchKernelGGL((k_c2r_rows<R, M, NL, T>), dim3((unsigned)nb), dim3(T), 0, st, in, out, nrows, (long long)P, tw,
                       sc, store, aux, part, (K > 0 && K <= M) ? K : M + 1);
    if (acc)
        hipLaunchKernelGGL(k_c2r_partsum, dim3(1), dim3(1024), 0, st, part, (int)nb, acc, aux ? 1 : 0);
    PYL_HIP(hipGetLastError());
    return OK;
}

template <typename R, int N> struct ColShape {
    static constexpr int NL = (N >= 1024 && sizeof(R) == 4) ? 8 : Geo<R>::NL;   // 1024-point lines: 8 per block
    // 16 elements per thread (the prefetch doubles the first stage's registers)
    static constexpr int T = clamp_i(N * NL / 16, 64, 512);
    static constexpr int LDS = (N * NL + N) * (int)sizeof(cplx<R>);
};

template <typename R, int N>
static int cols_job_n(const cplx<R>* src, cplx<R>* dst, int64_t P, int axis, R sc, int K, ColArgs<R>* a,
                      unsigned* blocks) {
    using SH = ColShape<R, N>;
    const cplx<R>* tw = twiddles<R>(N);
    PYL_CHECK(tw, E_ALLOC, "twiddle table allocation failed");
    // K < W: just the columns kz < K, enumerated without the others (column c = u * K + kz of the
    // outer coordinate u: x for the pass along y, y for the pass along x)
    const bool lim = K > 0 && K < N / 2 + 1;
    const long long PW = lim ? K : P;
    const int WV = lim ? K : N / 2 + 1;
    const long long ncols = (long long)N * PW;
    long long CW, CS, EST;
    if (axis == 1) {
        CW = PW;
        CS = (long long)N * P;
        EST = P;
    } else if (lim) {
        CW = PW;
        CS = P;
        EST = (long long)N * P;
    } else {
        CW = ncols;
        CS = 0;
        EST = (long long)N * P;
    }
    const long long nblk = (ncols + SH::NL - 1) / SH::NL;
    // persistent grid: about as many blocks as fit on the chip at once (LDS bound)
    int per_cu = (160 * 1024) / (SH::LDS + 1024);
    per_cu = per_cu < 1 ? 1 : (per_cu > 8 ? 8 : per_cu);
    long long nb = (long long)cu_count() * per_cu;
    if (nb > nblk) nb = nblk;
    *a = ColArgs<R>{src, dst, ncols, CW, CS, EST, PW, WV, tw, sc};
    *blocks = (unsigned)nb;
    return OK;
}

template <typename R, int N>
static int launch_cols(const cplx<R>* src, cplx<R>* dst, int64_t P, int axis, bool inv, R sc, hipStream_t st,
                       int K) {
    using SH = ColShape<R, N>;
    ColArgs<R> a;
    unsigned nb = 0;
    int rc = cols_job_n<R, N>(src, dst, P, axis, sc, K, &a, &nb);
    if (rc) return rc;
    if (inv)
        hipLaunchKernelGGL((k_cols<R, N, SH::NL, SH::T, true>), dim3(nb), dim3(SH::T), 0, st, a);
    else
        hipLaunchKernelGGL((k_cols<R, N, SH::NL, SH::T, false>), dim3(nb), dim3(SH::T), 0, st, a);
    PYL_HIP(hipGetLastError());
    return OK;
}

template <typename R, int N>
static int launch_cols_batch(const ColArgs<R>* tab, const unsigned* start, int nj, unsigned total, hipStream_t st) {
    using SH = ColShape<R, N>;
    hipLaunchKernelGGL((k_cols_b<R, N, SH::NL, SH::T, false>), dim3(total), dim3(SH::T), 0, st, tab, start, nj);
    PYL_HIP(hipGetLastError());
    return OK;
}

template <typename R>
int fft_pass_z_r2c(const R* in, cplx<R>* out, int N, int64_t P, bool affine, const double* dmass, double mass,
                   hipStream_t st) {
    const long long nrows = (long long)N * N;
    switch (N) {
#define PYL_CASE(n) \
    case n:         \
        return launch_r2c_rows<R, n>(in, out, nrows, P, affine, dmass, mass, st);
        PYL_FFT_SIZES(PYL_CASE)
#undef PYL_CASE
        default:
            set_error("FFT size %d not supported", N);
            return E_UNSUPPORTED;
    }
}

template <typename R>
int fft_pass_cols(const cplx<R>* src, cplx<R>* dst, int N, int64_t P, int axis, bool inverse, R sc,
                  hipStream_t st, int K) {
    switch (N) {
#define PYL_CASE(n) \
    case n:         \
        return launch_cols<R, n>(src, dst, P, axis, inverse, sc, st, K);
        PYL_FFT_SIZES(PYL_CASE)
#undef PYL_CASE
        default:
            set_error("FFT size %d not supported", N);
            return E_UNSUPPORTED;
    }
}

template <typename R>
int fft_pass_z_c2r(const cplx<R>* in, R* out, int N, int64_t P, R sc, int store, const R* aux, double* acc,
                   hipStream_t st, int K) {
    const long long nrows = (long long)N * N;
    switch (N) {
#define PYL_CASE(n) \
    case n:         \
        return launch_c2r_rows<R, n>(in, out, nrows, P, sc, store, aux, acc, st, K);
        PYL_FFT_SIZES(PYL_CASE)
#undef PYL_CASE
        default:
            set_error("FFT size %d not supported", N);
            return E_UNSUPPORTED;
    }
}

template <typename R>
int fft3d_r2c(const R* in, cplx<R>* out, int N, int64_t P, hipStream_t st, int K) {
    int rc = fft_pass_z_r2c<R>(in, out, N, P, false, nullptr, 1.0, st);
    if (rc) return rc;
    rc = fft_pass_cols<R>(out, out, N, P, 1, false, (R)1, st, K);
    if (rc) return rc;
    return fft_pass_cols<R>(out, out, N, P, 0, false, (R)1, st, K);
}

template <typename R>
int fft3d_c2r(const cplx<R>* in, R* out, cplx<R>* ws, int N, int64_t P, R sc, hipStream_t st) {
    int rc = fft_pass_cols<R>(in, ws, N, P, 0, true, (R)1, st);
    if (rc) return rc;
    rc = fft_pass_cols<R>(ws, ws, N, P, 1, true, (R)1, st);
    if (rc) return rc;
    return fft_pass_z_c2r<R>(ws, out, N, P, sc, 1, nullptr, nullptr, st);
}

template <typename R>
int r2c_job(const R* in, cplx<R>* out, int N, int64_t P, bool affine, const double* dmass, double mass,
            R2cArgs<R>* a, unsigned* blocks) {
    switch (N) {
#define PYL_CASE(n) \
    case n:         \
        return r2c_job_n<R, n>(in, out, (long long)N * N, P, affine, dmass, mass, a, blocks);
        PYL_FFT_SIZES(PYL_CASE)
#undef PYL_CASE
        default:
            set_error("FFT size %d not supported", N);
            return E_UNSUPPORTED;
    }
}

template <typename R>
int r2c_batch(int N, const R2cArgs<R>* tab, const unsigned* start, int nj, unsigned total, hipStream_t st) {
    switch (N) {
#define PYL_CASE(n) \
    case n:         \
        return launch_r2c_batch<R, n>(tab, start, nj, total, st);
        PYL_FFT_SIZES(PYL_CASE)
#undef PYL_CASE
        default:
            set_error("FFT size %d not supported", N);
            return E_UNSUPPORTED;
    }
}

template <typename R>
int cols_job(const cplx<R>* src, cplx<R>* dst, int N, int64_t P, int axis, ColArgs<R>* a, unsigned* blocks) {
    switch (N) {
#define PYL_CASE(n) \
    case n:         \
        return cols_job_n<R, n>(src, dst, P, axis, (R)1, 0, a, blocks);
        PYL_FFT_SIZES(PYL_CASE)
#undef PYL_CASE
        default:
            set_error("FFT size %d not supported", N);
            return E_UNSUPPORTED;
    }
}

template <typename R>
int cols_batch(int N, const ColArgs<R>* tab, const unsigned* start, int nj, unsigned total, hipStream_t st) {
    switch (N) {
#define PYL_CASE(n) \
    case n:         \
        return launch_cols_batch<R, n>(tab, start, nj, total, st);
        PYL_FFT_SIZES(PYL_CASE)
#undef PYL_CASE
        default:
            set_error("FFT size %d not supported", N);
            return E_UNSUPPORTED;
    }
}

#define PYL_INST(R)                                                                                        \
    template int fft_pass_z_r2c<R>(const R*, cplx<R>*, int, int64_t, bool, const double*, double, hipStream_t);               \
    template int fft_pass_cols<R>(const cplx<R>*, cplx<R>*, int, int64_t, int, bool, R, hipStream_t, int);    \
    template int fft_pass_z_c2r<R>(const cplx<R>*, R*, int, int64_t, R, int, const R*, double*, hipStream_t, int); \
    template int fft3d_r2c<R>(const R*, cplx<R>*, int, int64_t, hipStream_t, int);                                \
    template int fft3d_c2r<R>(const cplx<R>*, R*, cplx<R>*, int, int64_t, R, hipStream_t);                   \
    template int r2c_job<R>(const R*, cplx<R>*, int, int64_t, bool, const double*, double, R2cArgs<R>*, unsigned*); \
    template int r2c_batch<R>(int, const R2cArgs<R>*, const unsigned*, int, unsigned, hipStream_t);          \
    template int cols_job<R>(const cplx<R>*, cplx<R>*, int, int64_t, int, ColArgs<R>*, unsigned*);          \
    template int cols_batch<R>(int, const ColArgs<R>*, const unsigned*, int, unsigned, hipStream_t);
PYL_INST(float)
PYL_INST(double)
#undef PYL_INST

}  // namespace pyl
