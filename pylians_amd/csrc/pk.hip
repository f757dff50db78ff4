// Power spectrum on gfx950: bin geometry, the x-pass FFT fused with window deconvolution
// and k-shell binning, and the normalisation of the outputs.
//
// Semantics follow Pk_library.pyx:266-425:
//  * independent modes: kz in [0, N/2]; on the kz = 0 and kz = N/2 planes skip kx < 0, and where
//    kx is 0 or N/2 also skip ky < 0 (Pk_library.pyx:322-330);
//  * 3-D bin = floor(|k|) (integer k^2, exact isqrt), 1-D bin = |k_par| for |k| <= N/2,
//    2-D bin = (kmax_par+1) * floor(sqrt(k_perp^2)) + |k_par| (Pk_library.pyx:334-373);
//  * window factor prod_i (x_i/sin x_i)^p, x_i = pi k_i/N (Pk_library.pyx:75-87, 354-355);
//  * multipole weights L2 = (3mu^2-1)/2, L4 = (35mu^4-30mu^2+3)/8, mu = k_par/|k| (:346-348, :376-380);
//  * phase = atan2(Re, |delta|) (:361).
// Per-mode arithmetic is in the working precision R; every shell sum is float64.
#include <math.h>

#include <algorithm>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <type_traits>

#include "fft.hpp"
#include "fft_engine.hpp"
#include "pk.hpp"
#include "util.hpp"

namespace pyl {

static int isqrt_host(long long v) {
    long long r = (long long)sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return (int)r;
}

PkDims pk_dims(int N, int64_t P) {
    PkDims d;
    d.N = N;
    d.middle = N / 2;
    d.kmax_par = d.middle;
    d.kmax_per = isqrt_host(2LL * d.middle * d.middle);
    d.kmax = isqrt_host(3LL * d.middle * d.middle);
    d.P = P;
    return d;
}

size_t pk_out_len(const PkDims& d) { return (size_t)6 * d.kmax + 3 * (size_t)d.kmax_par + 4 * (size_t)d.n2(); }

__device__ __forceinline__ bool indep_mode(int kx, int ky, int kz, int middle, bool even) {
    if (kz == 0 || (kz == middle && even)) {
        if (kx < 0) return false;
        if ((kx == 0 || (kx == middle && even)) && ky < 0) return false;
    }
    return true;
}

// ------------------------------------------------------------------ geometry
// Counts and |k| sums of the independent modes, computed once per (N, axis) on the host with
// exact integer arithmetic; threads own slabs of kx and their histograms are summed in a fixed
// order, so the float64 |k| sums are reproducible.
static void geom_slab(int N, int axis, const PkDims& d, int x0, int x1, std::vector<unsigned long long>& n1,
                      std::vector<unsigned long long>& n2, std::vector<unsigned long long>& n3,
                      std::vector<double>& k3) {
    const int middle = N / 2;
    const bool even = (N % 2) == 0;
    for (int kxx = x0; kxx < x1; ++kxx) {
        const int kx = kxx > middle ? kxx - N : kxx;
        for (int kyy = 0; kyy < N; ++kyy) {
            const int ky = kyy > middle ? kyy - N : kyy;
            for (int kz = 0; kz <= middle; ++kz) {
                if (kz == 0 || (kz == middle && even)) {
                    if (kx < 0) continue;
                    if ((kx == 0 || (kx == middle && even)) && ky < 0) continue;
                }
                const long long k2 = (long long)kx * kx + (long long)ky * ky + (long long)kz * kz;
                const int bin = isqrt_host(k2);
                n3[bin] += 1;
                k3[bin] += sqrt((double)k2);
                int kpar, kper;
                if (axis == 0) {
                    kpar = kx;
                    kper = isqrt_host((long long)ky * ky + (long long)kz * kz);
                } else if (axis == 1) {
                    kpar = ky;
                    kper = isqrt_host((long long)kx * kx + (long long)kz * kz);
                } else {
                    kpar = kz;
                    kper = isqrt_host((long long)kx * kx + (long long)ky * ky);
                }
                kpar = kpar < 0 ? -kpar : kpar;
                if (k2 <= (long long)middle * middle) n1[kpar] += 1;
                n2[(size_t)(d.kmax_par + 1) * kper + kpar] += 1;
            }
        }
    }
}

// host half of the geometry (device pointers null), cached per (N, axis)
static const PkGeometry* host_geometry(int N, int axis) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, PkGeometry*> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(N, axis);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    PkGeometry* g = new PkGeometry();
    g->d = pk_dims(N, half_len(N));
    g->axis = axis;
    g->n1 = g->n2 = g->n3 = nullptr;
    g->k3sum = nullptr;
    const PkDims& d = g->d;
    g->h_n1.assign(d.n1(), 0);
    g->h_n2.assign(d.n2(), 0);
    g->h_n3.assign(d.n3(), 0);
    g->h_k3sum.assign(d.n3(), 0.0);
    unsigned nth = std::thread::hardware_concurrency();
    nth = nth < 1 ? 1 : (nth > 16 ? 16 : nth);
    if ((int)nth > N) nth = N;
    std::vector<std::vector<unsigned long long>> a1(nth, std::vector<unsigned long long>(d.n1(), 0)),
        a2(nth, std::vector<unsigned long long>(d.n2(), 0)), a3(nth, std::vector<unsigned long long>(d.n3(), 0));
    std::vector<std::vector<double>> ak(nth, std::vector<double>(d.n3(), 0.0));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; ++t) {
        const int x0 = (int)((long long)N * t / nth), x1 = (int)((long long)N * (t + 1) / nth);
        pool.emplace_back(geom_slab, N, axis, std::cref(d), x0, x1, std::ref(a1[t]), std::ref(a2[t]),
                          std::ref(a3[t]), std::ref(ak[t]));
    }
    for (auto& th : pool) th.join();
    for (unsigned t = 0; t < nth; ++t) {
        for (int i = 0; i < d.n1(); ++i) g->h_n1[i] += a1[t][i];
        for (int i = 0; i < d.n2(); ++i) g->h_n2[i] += a2[t][i];
        for (int i = 0; i < d.n3(); ++i) {
            g->h_n3[i] += a3[t][i];
            g->h_k3sum[i] += ak[t][i];
        }
    }
    cache[key] = g;
    return g;
}

// Device geometry: cached per (device, N, axis); inside a fixed arena, a per-thread copy whose
// device arrays are fresh uploads into the arena.
const PkGeometry* pk_geometry(int N, int axis, int* rc) {
    *rc = OK;
    const PkGeometry* hg = host_geometry(N, axis);
    const PkDims& d = hg->d;
    const size_t b1 = sizeof(unsigned long long) * d.n1(), b2 = sizeof(unsigned long long) * d.n2(),
                 b3 = sizeof(unsigned long long) * d.n3(), bk = sizeof(double) * d.n3();
    if (arena_fixed()) {
        static thread_local PkGeometry tg;
        tg = *hg;
        tg.n1 = (unsigned long long*)arena_upload(hg->h_n1.data(), b1, rc);
        if (*rc) return nullptr;
        tg.n2 = (unsigned long long*)arena_upload(hg->h_n2.data(), b2, rc);
        if (*rc) return nullptr;
        tg.n3 = (unsigned long long*)arena_upload(hg->h_n3.data(), b3, rc);
        if (*rc) return nullptr;
        tg.k3sum = (double*)arena_upload(hg->h_k3sum.data(), bk, rc);
        if (*rc) return nullptr;
        return &tg;
    }
    static std::mutex mu;
    static std::map<std::tuple<int, int, int>, PkGeometry*> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("hipGetDevice failed");
        *rc = E_LAUNCH;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_tuple(dev, N, axis);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    PkGeometry* g = new PkGeometry(*hg);
    char* base = nullptr;
    if (hipMalloc((void**)&base, b1 + b2 + b3 + bk) != hipSuccess) {
        set_error("geometry allocation failed");
        *rc = E_ALLOC;
        delete g;
        return nullptr;
    }
    g->k3sum = (double*)base;
    g->n1 = (unsigned long long*)(base + bk);
    g->n2 = (unsigned long long*)(base + bk + b1);
    g->n3 = (unsigned long long*)(base + bk + b1 + b2);
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMemcpy(g->n1, g->h_n1.data(), b1, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g->n2, g->h_n2.data(), b2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g->n3, g->h_n3.data(), b3, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g->k3sum, g->h_k3sum.data(), bk, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error("geometry upload failed: %s", hipGetErrorString(e));
        *rc = E_LAUNCH;
        (void)hipFree(base);
        delete g;
        return nullptr;
    }
    cache[key] = g;
    return g;
}

// ------------------------------------------------------------------ window tables
// win[i] = (x/sin x)^p with x = pi kfreq(i)/N (1 at x = 0): Pk_library.pyx:86-87, 316-324
static const std::vector<double>& host_window(int N, int p) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::vector<double>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(N, p);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    std::vector<double>& h = cache[key];
    h.resize(N);
    for (int i = 0; i < N; ++i) {
        int k = i > N / 2 ? i - N : i;
        double x = M_PI / N * k;
        h[i] = (x == 0.0) ? 1.0 : pow(x / sin(x), p);
    }
    return h;
}

// device copy cached per (device, N, p); inside a fixed arena a fresh copy in the arena
static const double* window_table(int N, int p) {
    const std::vector<double>& h = host_window(N, p);
    if (arena_fixed()) {
        int rc = OK;
        return (const double*)arena_upload(h.data(), sizeof(double) * N, &rc);
    }
    static std::mutex mu;
    static std::map<std::tuple<int, int, int>, double*> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_tuple(dev, N, p);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    double* dptr = nullptr;
    if (hipMalloc((void**)&dptr, sizeof(double) * N) != hipSuccess) return nullptr;
    if (hipMemcpy(dptr, h.data(), sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    cache[key] = dptr;
    return dptr;
}
const double* pk_window_table(int N, int p) { return window_table(N, p); }

// ------------------------------------------------------------------ last pass + binning
constexpr int isqrt_c(long long v) {
    long long r = 0;
    while ((r + 1) * (r + 1) <= v) ++r;
    return (int)r;
}
// LDS bytes of the histograms: 3-D (4 float64 sums + a mode count), 1-D and the 2-D window (one
// k_par row; NL rows when blocks own a kz block, fixk = 1), each a float64 sum and a 32-bit count
constexpr int hist_bytes(int N, int NL, bool fixk) {
    return (isqrt_c(3LL * (N / 2) * (N / 2)) + 1) * (4 * 8 + 4) + (N / 2 + 1) * 12 +
           (fixk ? NL : 1) * (isqrt_c(2LL * (N / 2) * (N / 2)) + 1) * 12;
}
// Threads of the binning kernel: N/16 per column line, doubled (measured faster) for lines with a
// radix-3 factor from 192 points up and where bin_wide says so: blocks that own a kz block (line of
// sight z) hold a 2-D histogram row per line, which leaves room for few blocks per CU, and from 256
// points up twice the waves per block paid (0.73 -> 0.55 ms at 512); with the line of sight along x
// or y, 256 points gained (0.094 -> 0.085 ms) and 512 lost (0.397 -> 0.410 ms).
constexpr bool bin_wide(int N, bool fixk) { return fixk ? N >= 256 : (N >= 64 && N <= 256); }
constexpr int threads_bin_c(int N, int NL, bool fixk) {
    return clamp_i(N * NL / 16 * ((N % 3 == 0 && N >= 192) || bin_wide(N, fixk) ? 2 : 1), 64, 512);
}
// LDS rows of the binning kernel's transform (RowMap::ROWS for its EU), at least N + T / NL
// (rows of NL + 1 values in the mapped layout)
constexpr int bin_rows(int N, int T, int NL) {
    const int eu = (N / 2 + 1) / (T / NL) > 0 ? (N / 2 + 1) / (T / NL) : 1;
    const int rows = eu % 2 == 0 ? N + 2 * ((N / 2) / eu) + 2 : N;
    return rows > N + T / NL ? rows : N + T / NL;
}
// LDS bytes of one binning block: the mapped rows, the twiddle and window tables, the histograms
template <typename R> constexpr int lds_bin(int N, int T, int NL, bool fixk) {
    return bin_rows(N, T, NL) * (NL + 1) * (int)sizeof(cplx<R>) + N * (int)(sizeof(cplx<R>) + sizeof(R)) +
           hist_bytes(N, NL, fixk);
}
// lines per block when two blocks share a CU (160 KB of LDS) beside their histograms: at most 8
template <typename R, int N, bool FIXK> constexpr int nl_pair() {
    int nl = 8;
    while (nl > 2 && lds_bin<R>(N, threads_bin_c(N, nl, FIXK), nl, FIXK) > 80 * 1024) nl /= 2;
    return nl;
}
// Where two blocks per CU leave only 2 lines per block (float64 lines from 512 points, float32 from
// 640: 32- or 16-byte column segments and 1-2 waves per block), one block per CU with 8 lines, or 4
// where 8 do not fit, and N * lines / 8 threads (at most 512) (float64 at 512: line of sight x
// 1.42 -> 0.69 ms with 8 lines, z 1.27 -> 0.99 ms with 4)
constexpr int SOLO_LDS = 150 * 1024;
constexpr int solo_threads(int N, int nl) { return clamp_i(N * nl / 8, 64, 512); }
template <typename R, int N, bool FIXK> constexpr int solo_nl() {
    if (nl_pair<R, N, FIXK>() != 2) return 0;
    if (lds_bin<R>(N, solo_threads(N, 8), 8, FIXK) <= SOLO_LDS) return 8;
    if (lds_bin<R>(N, solo_threads(N, 4), 4, FIXK) <= SOLO_LDS) return 4;
    return 0;
}
template <typename R, int N, bool FIXK> constexpr int nl_bin() {
    return solo_nl<R, N, FIXK>() ? solo_nl<R, N, FIXK>() : nl_pair<R, N, FIXK>();
}
template <typename R, int N, bool FIXK> constexpr int threads_bin() {
    return solo_nl<R, N, FIXK>() ? solo_threads(N, solo_nl<R, N, FIXK>())
                                 : threads_bin_c(N, nl_bin<R, N, FIXK>(), FIXK);
}


// Squared phase atan2(re, |v|)^2 of a mode (Pk_library.pyx:374), r2 = |v|^2 >= re^2.  float32:
// t = |re|/|v| in [0, 1] and atan(t) = t P(t^2) with a degree-7 polynomial fitted (Lawson-weighted
// least squares) to atan(t)/t on the whole of [0, 1], so there is no range reduction and no branch:
// 2.1e-7 relative error in float32.  r2 = 0 (then re = 0) gives 0.  float64: the library atan2.
__device__ __forceinline__ float phase_sq(float re, float r2) {
    // t may exceed 1 by the rounding of rsq; the polynomial holds there too
    const float t = fabsf(re) * __builtin_amdgcn_rsqf(r2 + 1.17549435e-38f);
    const float s = t * t;
    float p = -0.004693399649f;
    p = p * s + 0.024252869189f;
    p = p * s - 0.059487096965f;
    p = p * s + 0.099143467844f;
    p = p * s - 0.140195026994f;
    p = p * s + 0.199697285891f;
    p = p * s - 0.333319902420f;
    p = p * s + 0.999999880791f;
    const float a = t * p;
    return a * a;
}
__device__ __forceinline__ double phase_sq(double re, double r2) {
    const double ph = atan2(re, sqrt(r2));
    return ph * ph;
}

// LDS row of FFT output k in the binning kernel.  The epilogue pairs rows k and N-k (same |k|);
// thread jre of a column owns the EU pairs from u = jre*EU (the last few threads one more, each
// shifted by one per thread before it that took one more).  Rows hold MP = NL + 1 values (8-byte
// units), so the rows that consecutive threads read at one step are an odd number of units apart
// (EU rows, plus one spare row per group of EU rows when EU is even) and fall on distinct LDS banks.
template <int N, int EU> struct RowMap {
    static constexpr int PADR = (EU % 2 == 0) ? 1 : 0;
    static constexpr int GL = (N / 2) / EU + 1;   // groups of the lower half k <= N/2
    static constexpr int GU = (N / 2) / EU;
    static constexpr int ROWS = N + PADR * (GL + GU + 1);
    static __device__ __forceinline__ int row(int k) {
        if (PADR == 0) return k;
        if (k <= N / 2) return k + k / EU;
        return k + GL + GU - (N - k) / EU;
    }
};

// Last radix stage of the forward transform, written back to LDS in the RowMap order with MP
// values per row.
template <typename R, int N, int NL, int T, class RM, int MP>
__device__ __forceinline__ void last_stage_mapped(cplx<R>* s, const cplx<R>* tw, int tid) {
    constexpr Radices PL = make_radices(N);
    constexpr int RAD = PL.r[PL.n - 1];
    constexpr int NS = N / RAD;
    constexpr int NBF = NS * NL;
    constexpr int B = (NBF + T - 1) / T;
    cplx<R> v[B][RAD];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int g = tid + b * T;
        if (NBF % T == 0 || g < NBF) {
            const int l = g % NL, j = g / NL;
#pragma unroll
            for (int m = 0; m < RAD; ++m) v[b][m] = s[(j + m * NS) * NL + l];
            if (NS > 1) {
#pragma unroll
                for (int m = 1; m < RAD; ++m) v[b][m] = cmul(v[b][m], tw[m * j]);
            }
            Dft<R, RAD>::run(v[b]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int g = tid + b * T;
        if (NBF % T == 0 || g < NBF) {
            const int l = g % NL, j = g / NL;
#pragma unroll
            for (int m = 0; m < RAD; ++m) s[RM::row(j + m * NS) * MP + l] = v[b][m];
        }
    }
    __syncthreads();
}

// The last FFT pass of the forward transform, fused with the window deconvolution and the
// binning.  The pass runs along x (last = 0; column (u = y, kz)) or y (last = 1; column
// (u = x, kz)), chosen so that k_par is constant along each column (line of sight x or y: k_par =
// k_u; z: k_par = kz).  Blocks own either one u and a range of kz blocks ("fixk = 0") or one kz
// block and a range of u ("fixk = 1"), so their 2-D histogram window stays in LDS.  A block walks
// its column blocks in order; the first-stage inputs of the next one are loaded into registers
// while the current one is transformed and binned.  The spectrum has row pitch P (kz < W valid);
// the optional window-corrected copy `keep` has pitch W.
// Every histogram entry also counts the modes it received (Nmodes, Pk_library.pyx:376-381): a
// block's row of `part` is [3-D sums (4 n3) | 1-D sums (n1) | 3-D counts (n3) | 1-D counts (n1) |
// 2-D sums (n2w) | 2-D counts (n2w)], counts held as float64.

template <typename R, int N, int NL, int T>
__device__ __forceinline__ void last_bin_body(const LastBinArgs<R>& A, Blk B) {
    const cplx<R>* __restrict__ src = A.src;
    cplx<R>* __restrict__ keep = A.keep;
    const long long P = A.P;
    const int W = A.W, last = A.last, axis = A.axis, fixk = A.fixk, nchunk = A.nchunk, flags = A.flags;
    const int kmax_par = A.kmax_par, kmax_per = A.kmax_per, n1 = A.n1, n3 = A.n3;
    const double* __restrict__ wing = A.win;
    const cplx<R>* __restrict__ twg = A.tw;
    double* __restrict__ part = A.part;
    const int RS = A.RS;
    constexpr int TPC = T / NL;              // threads per column
    constexpr int U = N / 2 + 1;             // row pairs (k, N-k), k = 0 .. N/2 (k = 0, N/2 alone)
    constexpr int EU = U / TPC;              // consecutive pairs per thread; the last UR threads of a
    constexpr int UR = U % TPC;              // column take one more
    static_assert(T % NL == 0, "bad binning geometry");
    using F = First<N, NL, T>;
    using RM = RowMap<N, (EU > 0 ? EU : 1)>;
    constexpr int MP = NL + 1;               // values per row of the mapped layout
    constexpr bool ONE = F::NST == 1;
    __shared__ cplx<R> s[(RM::ROWS * MP > N * NL ? RM::ROWS * MP : N * NL)];
    __shared__ cplx<R> tw[N];
    __shared__ R win[N];
    extern __shared__ double hs[];
    const int nper = kmax_per + 1;
    const int n2w = (fixk ? NL : 1) * nper;
    const int nsum = 4 * n3 + n1 + n2w;
    double* h3 = hs;             // n3 * 4
    double* h1 = hs + 4 * n3;    // n1
    double* h2 = h1 + n1;        // NL * nper (fixk = 1) or nper (fixk = 0)
    unsigned* c3 = (unsigned*)(hs + nsum);   // mode counts: n3 | n1 | n2w
    unsigned* c1 = c3 + n3;
    unsigned* c2 = c1 + n1;
    const int tid = threadIdx.x;
    for (int i = tid; i < N; i += T) {
        tw[i] = twg[i];
        win[i] = (R)wing[i];
    }
    for (int i = tid; i < nsum; i += T) hs[i] = 0.0;
    for (int i = tid; i < n3 + n1 + n2w; i += T) c3[i] = 0u;
    const bool drop0 = flags & PK_DROP_KZ0;

    const int middle = N / 2;
    const bool even = (N % 2) == 0;
    const int nkzb = (W + NL - 1) / NL;
    const long long EST = last == 0 ? (long long)N * P : P;   // element stride along the pass axis
    const long long UST = last == 0 ? P : (long long)N * P;   // stride of the outer coordinate u
    const long long KEST = last == 0 ? (long long)N * W : W;  // the same in `keep`
    const long long KUST = last == 0 ? W : (long long)N * W;
    const int l = tid % NL;
    const bool do1 = flags & PK_1D, do2 = flags & PK_2D, doph = flags & PK_PHASE;
    // column blocks of this block: items it = 0 .. nit-1
    int u0, kzb0, nit;
    if (fixk) {
        const int kzb = B.x / nchunk, ch = B.x % nchunk;
        const int UC = (N + nchunk - 1) / nchunk;
        u0 = ch * UC;
        nit = max(0, min(N, u0 + UC) - u0);
        kzb0 = kzb;
    } else {
        const int KB = (nkzb + nchunk - 1) / nchunk;
        u0 = B.x / nchunk;
        kzb0 = (B.x % nchunk) * KB;
        nit = max(0, min(nkzb, kzb0 + KB) - kzb0);
    }
    auto item = [&](int it, int& u, int& kzz) {
        u = fixk ? u0 + it : u0;
        kzz = (fixk ? kzb0 : kzb0 + it) * NL + l;
    };
    auto fetch = [&](int it, cplx<R> (&v)[F::B][F::RAD]) {
        int u, kzz;
        item(it, u, kzz);
        const bool ok = kzz < W;
        const long long b = (long long)u * UST + kzz;
        first_load<R, N, NL, T>(v, [&](int j, int) { return ok ? src[b + j * EST] : mk<R>(0, 0); }, tid);
    };
    cplx<R> nxt[F::B][F::RAD];
    if (nit > 0) fetch(0, nxt);
    for (int it = 0; it < nit; ++it) {
        {
            cplx<R> cur[F::B][F::RAD];
#pragma unroll
            for (int b = 0; b < F::B; ++b)
#pragma unroll
                for (int m = 0; m < F::RAD; ++m) cur[b][m] = nxt[b][m];
            if (it + 1 < nit) fetch(it + 1, nxt);
            __syncthreads();   // the previous epilogue has read s
            if constexpr (ONE) {
                // a single radix stage: straight from registers to the padded rows
#pragma unroll
                for (int b = 0; b < F::B; ++b) {
                    const int g = tid + b * T;
                    if (F::NBF % T == 0 || g < F::NBF) {
                        const int jj = g / NL;
                        Dft<R, F::RAD>::run(cur[b]);
#pragma unroll
                        for (int m = 0; m < F::RAD; ++m) s[RM::row(jj * F::RAD + m) * MP + l] = cur[b][m];
                    }
                }
                __syncthreads();
            } else {
                first_stage_regs<R, N, NL, NL, T>(s, cur, tid);
            }
        }
        if constexpr (!ONE) {
            mid_stages<R, N, NL, NL, T, 1, 1, F::RAD>(s, tw, tid);
            last_stage_mapped<R, N, NL, T, RM, MP>(s, tw, tid);
        }
        int u, kzz;
        item(it, u, kzz);
        // the first kz of the column block (the same for the whole block): only the blocks that hold
        // kz = 0 or kz = N/2 apply the independent-mode rule of those planes (Pk_library.pyx:322-330)
        const int kzlo = kzz - l;
        // the epilogue takes its own lane layout: TPC consecutive threads share a column (le) and hold
        // consecutive runs of row pairs (jre), so that the threads of a wave flush different bins
        // (with consecutive columns per wave they mostly share one, and their LDS atomics serialise)
        const int le = tid / TPC, jre = tid % TPC;
        kzz = kzlo + le;
        const bool zblock = kzlo == 0 || (even && middle >= kzlo && middle < kzlo + NL);
        if (kzz >= W || (drop0 && kzz == 0)) continue;
        const int kz = kzz;
        const int ku = kfreq(u, N);
        const R cuz = win[u] * win[kzz];
        const bool zplane = (kz == 0) || (kz == middle && even);
        // k_par is constant on the column: k_u (line of sight x or y) or kz
        const int kps = axis == 2 ? kz : ku;
        const int c0 = ku * ku + kz * kz;
        const int cper = c0 - kps * kps;
        const R kps2 = (R)(kps * kps);
        const long long kbase = (long long)u * KUST + kzz;
        // |k|^2 = c0 + u^2 and k_perp^2 = cper + u^2 of this thread's first pair and their bins
        // floor(sqrt); each step of u adds at most 1 to sqrt(c + u^2) (its derivative u / sqrt(c + u^2)
        // is <= 1), so each bin grows by 0 or 1 per step and is carried along without a square root
        const int extra = jre - (TPC - UR);   // >= 0: this thread takes EU + 1 pairs
        const int uq0 = jre * EU + (extra > 0 ? extra : 0);
        const int ne = EU + (extra >= 0 ? 1 : 0);
        int k2 = c0 + uq0 * uq0, kp2 = cper + uq0 * uq0;
        int bin = isqrt24(k2), wb = isqrt24(kp2);
        const int w2o = fixk ? le * nper : 0;
        // the current runs: 3-D bin cur3 and 2-D entry cur2 with their sums and modes; a run is added
        // to the LDS histograms when the bin changes and at the end (if it holds modes)
        int cur3 = bin, cur2 = w2o + wb;
        double a0 = 0.0, a2 = 0.0, a4 = 0.0, ap = 0.0, b1 = 0.0, b2 = 0.0;
        unsigned m3 = 0u, m1 = 0u, m2 = 0u;
        // pairs of rows with opposite k_j: the same |k|, k_perp, mu^2 and window factor.  ZP: the
        // column block holds a kz = 0 or Nyquist column, where a mode may not be independent
        auto pairs = [&](auto zp) {
            constexpr bool ZP = decltype(zp)::value;
#pragma unroll 1
            for (int q = 0; q < ne; ++q) {
                const int uu = uq0 + q;
                if (q > 0) {
                    k2 += 2 * uu - 1;
                    kp2 += 2 * uu - 1;
                    bin += __mul24(bin + 1, bin + 1) <= k2 ? 1 : 0;
                    wb += __mul24(wb + 1, wb + 1) <= kp2 ? 1 : 0;
                }
                const int ja = uu, jb = (N - uu) % N;
                const bool pair = jb != ja;
                const cplx<R> va = s[RM::row(ja) * MP + le];
                const cplx<R> vb = pair ? s[RM::row(jb) * MP + le] : mk<R>(0, 0);
                bool ia = true, ib = pair;
                if (ZP && zplane) {
                    const int xa = last == 0 ? uu : ku, ya = last == 0 ? ku : uu;
                    const int xb = last == 0 ? -uu : ku, yb = last == 0 ? ku : -uu;
                    // k_j = N/2 is +N/2 (kfreq); -uu for the partner row N - uu
                    ia = !(xa < 0 || ((xa == 0 || (xa == middle && even)) && ya < 0));
                    ib = pair && !(xb < 0 || ((xb == 0 || (xb == middle && even)) && yb < 0));
                }
                const R fac = win[uu] * cuz;
                if (keep) {
                    keep[kbase + ja * KEST] = ia ? scale(va, fac) : va;
                    if (pair) keep[kbase + jb * KEST] = ib ? scale(vb, fac) : vb;
                }
                if (ZP && !ia && !ib) continue;
                const unsigned nm = (ia ? 1u : 0u) + (ib ? 1u : 0u);
                const R ra = va.x * va.x + va.y * va.y;
                const R rb = vb.x * vb.x + vb.y * vb.y;
                const R r2 = (ia ? ra : (R)0) + (ib ? rb : (R)0);
                const R mu2 = k2 == 0 ? (R)0 : kps2 * recip((R)k2);
                const R d2 = r2 * (fac * fac);
                const R L2 = ((R)1.5 * mu2 - (R)0.5);
                const R L4 = (((R)35 * mu2 - (R)30) * mu2 + (R)3) * (R)0.125;
                if (bin != cur3) {
                    atomicAdd(&h3[4 * cur3 + 0], a0);
                    atomicAdd(&h3[4 * cur3 + 1], a2);
                    atomicAdd(&h3[4 * cur3 + 2], a4);
                    if (doph) atomicAdd(&h3[4 * cur3 + 3], ap);
                    atomicAdd(&c3[cur3], m3);
                    cur3 = bin;
                    a0 = a2 = a4 = ap = 0.0;
                    m3 = 0u;
                }
                const double dd = (double)d2;
                m3 += nm;
                a0 += dd;
                a2 += (double)(d2 * L2);
                a4 += (double)(d2 * L4);
                if (doph) {
                    // the phase of a mode is unchanged by the (positive) window factor
                    const R pa = phase_sq(va.x, ra), pb = phase_sq(vb.x, rb);
                    ap += (double)((ia ? pa : (R)0) + (ib ? pb : (R)0));
                }
                if (do1 && k2 <= middle * middle) {
                    b1 += dd;
                    m1 += nm;
                }
                if (do2) {
                    const int w2 = w2o + wb;
                    if (w2 != cur2) {
                        atomicAdd(&h2[cur2], b2);
                        atomicAdd(&c2[cur2], m2);
                        cur2 = w2;
                        b2 = 0.0;
                        m2 = 0u;
                    }
                    b2 += dd;
                    m2 += nm;
                }
            }
        };
        if (zblock)
            pairs(std::true_type{});
        else
            pairs(std::false_type{});
        if (m3 != 0u) {
            atomicAdd(&h3[4 * cur3 + 0], a0);
            atomicAdd(&h3[4 * cur3 + 1], a2);
            atomicAdd(&h3[4 * cur3 + 2], a4);
            if (doph) atomicAdd(&h3[4 * cur3 + 3], ap);
            atomicAdd(&c3[cur3], m3);
        }
        if (m1 != 0u) {
            atomicAdd(&h1[abs(kps)], b1);
            atomicAdd(&c1[abs(kps)], m1);
        }
        if (m2 != 0u) {
            atomicAdd(&h2[cur2], b2);
            atomicAdd(&c2[cur2], m2);
        }
    }
    // this block's histograms -> its row of the partials (summed by k_bin_reduce31 / k_bin_reduce2)
    __syncthreads();
    double* row = part + (long long)B.x * RS;
    const int n31 = 5 * n3 + 2 * n1;
    for (int i = tid; i < 4 * n3 + n1; i += T) row[i] = hs[i];
    for (int i = tid; i < n3 + n1; i += T) row[4 * n3 + n1 + i] = (double)c3[i];
    for (int i = tid; i < n2w; i += T) {
        row[n31 + i] = h2[i];
        row[n31 + n2w + i] = (double)c2[i];
    }
}
template <typename R, int N, int NL, int T>
__global__ __launch_bounds__(T) void k_last_bin(LastBinArgs<R> a) {
    last_bin_body<R, N, NL, T>(a, own_blk());
}
template <typename R, int N, int NL, int T>
__global__ __launch_bounds__(T) void k_last_bin_b(const LastBinArgs<R>* __restrict__ tab,
                                                   const unsigned* __restrict__ start, int nj) {
    PYL_BATCH_BLK(start, nj);
    last_bin_body<R, N, NL, T>(tab[bj_], blk_);
}

// Sum the per-block histogram rows of k_last_bin into the pass sums.
// 3-D and 1-D entries (k_bin_reduce31): 64 entries per workgroup, the blocks split over the 4 waves
// and over gridDim.y slices; each slice adds its total with one atomic.  Entries: 3-D sums, 1-D
// sums, 3-D counts, 1-D counts.

__device__ __forceinline__ void reduce31_body(const ReduceArgs& A) {
    const double* __restrict__ part = A.part;
    const int RS = A.RS, nblocks = A.nblocks, n3 = A.n3, n1 = A.n1;
    double* __restrict__ g3 = A.g3;
    double* __restrict__ g1 = A.g1;
    double* __restrict__ gc3 = A.gc3;
    double* __restrict__ gc1 = A.gc1;
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    const int n31 = 5 * n3 + 2 * n1;
    const int step = 4 * gridDim.y;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < n31) {
        int b = blockIdx.y * 4 + w, q = 0;
        for (; b < nblocks; b += step, q = (q + 1) & 3) a[q] += part[(long long)b * RS + i];
    }
    red[w][lane] = (a[0] + a[1]) + (a[2] + a[3]);
    __syncthreads();
    if (w == 0 && i < n31) {
        const double t = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
        double* dst = i < 4 * n3 ? &g3[i]
                      : i < 4 * n3 + n1 ? &g1[i - 4 * n3]
                      : i < 5 * n3 + n1 ? &gc3[i - 4 * n3 - n1]
                                        : &gc1[i - 5 * n3 - n1];
        if (t != 0.0) atomicAdd(dst, t);
    }
}
__global__ __launch_bounds__(256) void k_bin_reduce31(ReduceArgs a) { reduce31_body(a); }
// batched: job blockIdx.z
__global__ __launch_bounds__(256) void k_bin_reduce31_b(const ReduceArgs* __restrict__ tab) {
    const ReduceArgs& a = tab[blockIdx.z];
    reduce31_body(a);
}

// 2-D entries (k_bin_reduce2): (k_per, k_par) comes from the blocks whose window holds k_par:
// fixk = 0, the blocks of u = k_par and u = N - k_par (k_par = |k_u|); fixk = 1, the blocks of kz
// block k_par / NL, row k_par % NL.
// The 2-D counts sit n2w entries after the 2-D sums of a row.
__device__ __forceinline__ void reduce2_body(const ReduceArgs& A) {
    const double* __restrict__ part = A.part;
    const int RS = A.RS, nblocks = A.nblocks, n31 = 5 * A.n3 + 2 * A.n1, n2w = A.n2w, nper = A.nper;
    const int kmax_par = A.kmax_par, fixk = A.fixk, nchunk = A.nchunk, NL = A.NL, N = A.N;
    double* __restrict__ g2 = A.g2;
    double* __restrict__ gc2 = A.gc2;
    const long long i2 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i2 >= (long long)(kmax_par + 1) * nper) return;
    const int kpar = (int)(i2 % (kmax_par + 1)), kper = (int)(i2 / (kmax_par + 1));
    double a = 0.0, m = 0.0;
    if (fixk) {
        const int kzb = kpar / NL, r = kpar % NL;
        for (int c = 0; c < nchunk; ++c) {
            const long long b = (long long)kzb * nchunk + c;
            if (b < nblocks) {
                a += part[b * RS + n31 + r * nper + kper];
                m += part[b * RS + n31 + n2w + r * nper + kper];
            }
        }
    } else {
        const int us[2] = {kpar, N - kpar};
        const int nu = (kpar > 0 && N - kpar != kpar && N - kpar > N / 2) ? 2 : 1;
        for (int q = 0; q < nu; ++q)
            for (int c = 0; c < nchunk; ++c) {
                const long long b = (long long)us[q] * nchunk + c;
                a += part[b * RS + n31 + kper];
                m += part[b * RS + n31 + n2w + kper];
            }
    }
    if (a != 0.0) g2[i2] += a;
    if (m != 0.0) gc2[i2] += m;
}
__global__ void k_bin_reduce2(ReduceArgs a) { reduce2_body(a); }
// batched: job blockIdx.y
__global__ void k_bin_reduce2_b(const ReduceArgs* __restrict__ tab) {
    const ReduceArgs& a = tab[blockIdx.y];
    reduce2_body(a);
}

static int num_cus() {
    static int n = 0;
    if (!n) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
    }
    return n;
}

int pk_last_axis(int axis) { return axis == 0 ? 1 : 0; }

// launch shape of the binning pass for one grid size and line-of-sight class: blocks own one kz block
// (fixk, line of sight z) or one u; about two blocks per CU, all resident at once, each walking its
// column blocks with prefetch (four per CU with fixk measured faster)
template <typename R, int N, bool FIXK> struct BinShape {
    static constexpr int NL = nl_bin<R, N, FIXK>();
    static constexpr int T = threads_bin<R, N, FIXK>();
    static constexpr int W = N / 2 + 1;   // kz = 0 .. N/2; rows of the spectrum may be padded to P >= W
    int nchunk, grid, n2w, RS, n31;
    long long n2;
    size_t shm;
    PkDims d;
    BinShape() {
        d = pk_dims(N, W);
        const int nkzb = (W + NL - 1) / NL;
        const int want = (FIXK ? 4 : 2) * num_cus();
        if (FIXK) {
            nchunk = std::min(N, std::max(1, want / nkzb));
            grid = nkzb * nchunk;
        } else {
            nchunk = std::min(nkzb, std::max(1, (want + N - 1) / N));   // at least `want` blocks when it can
            grid = N * nchunk;
        }
        const int nper = d.kmax_per + 1;
        n2w = (FIXK ? NL : 1) * nper;
        RS = 5 * d.n3() + 2 * d.n1() + 2 * n2w;
        n31 = 5 * d.n3() + 2 * d.n1();
        n2 = (long long)(d.kmax_par + 1) * nper;
        // LDS: float64 sums (4 n3 + n1 + n2w), then 32-bit counts (n3 + n1 + n2w)
        shm = sizeof(double) * (size_t)(4 * d.n3() + d.n1() + n2w) + sizeof(unsigned) * (size_t)(d.n3() + d.n1() + n2w);
    }
    int slices() const { return std::max(1, std::min(16, grid / 32)); }
};

template <typename R, int N, bool FIXK>
static int last_bin_job_k(const cplx<R>* spec, cplx<R>* keep, int64_t P, int axis, int p, int flags,
                          const PkSums& sums, double* part, LastBinArgs<R>* la, ReduceArgs* ra, unsigned* blocks) {
    const BinShape<R, N, FIXK> sh;
    const cplx<R>* tw = twiddles<R>(N);
    const double* win = window_table(N, p);
    PYL_CHECK(tw && win, E_ALLOC, "table allocation failed");
    const PkDims& d = sh.d;
    const int fixk = FIXK ? 1 : 0;
    *la = LastBinArgs<R>{spec, keep, (long long)P, sh.W, pk_last_axis(axis), axis, fixk, sh.nchunk, flags,
                         d.kmax_par, d.kmax_per, d.n1(), d.n3(), win, tw, part, sh.RS};
    *ra = ReduceArgs{part, sh.RS, sh.grid, d.n3(), d.n1(), sh.n2w, d.kmax_per + 1, d.kmax_par, fixk, sh.nchunk,
                     sh.NL, N, sums.s3, sums.s1, sums.c3, sums.c1, sums.s2, sums.c2};
    *blocks = (unsigned)sh.grid;
    return OK;
}

template <typename R, int N, bool FIXK>
static int launch_last_bin_k(const cplx<R>* spec, cplx<R>* keep, int64_t P, int axis, int p, int flags,
                             const PkSums& sums, hipStream_t st) {
    const BinShape<R, N, FIXK> sh;
    int rc = OK;
    double* part = (double*)workspace(WS_PART, sizeof(double) * (size_t)sh.grid * sh.RS, &rc);
    if (rc) return rc;
    LastBinArgs<R> la;
    ReduceArgs ra;
    unsigned grid = 0;
    rc = last_bin_job_k<R, N, FIXK>(spec, keep, P, axis, p, flags, sums, part, &la, &ra, &grid);
    if (rc) return rc;
    hipLaunchKernelGGL((k_last_bin<R, N, sh.NL, sh.T>), dim3(grid), dim3(sh.T), sh.shm, st, la);
    PYL_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bin_reduce31, dim3((unsigned)((sh.n31 + 63) / 64), (unsigned)sh.slices()), dim3(256), 0, st,
                       ra);
    PYL_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bin_reduce2, dim3((unsigned)((sh.n2 + 255) / 256)), dim3(256), 0, st, ra);
    PYL_HIP(hipGetLastError());
    return OK;
}

template <typename R, int N, bool FIXK>
static int launch_last_bin_batch_k(const LastBinArgs<R>* tab, const unsigned* start, int nj, unsigned total,
                                   const ReduceArgs* rtab, hipStream_t st) {
    const BinShape<R, N, FIXK> sh;
    hipLaunchKernelGGL((k_last_bin_b<R, N, sh.NL, sh.T>), dim3(total), dim3(sh.T), sh.shm, st, tab, start, nj);
    PYL_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bin_reduce31_b, dim3((unsigned)((sh.n31 + 63) / 64), (unsigned)sh.slices(), (unsigned)nj),
                       dim3(256), 0, st, rtab);
    PYL_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bin_reduce2_b, dim3((unsigned)((sh.n2 + 255) / 256), (unsigned)nj), dim3(256), 0, st, rtab);
    PYL_HIP(hipGetLastError());
    return OK;
}

// blocks own one kz block (fixk) when the line of sight is z, else one u: separate LDS budgets
template <typename R, int N>
static int launch_last_bin(const cplx<R>* spec, cplx<R>* keep, int64_t P, int axis, int p, int flags,
                           const PkSums& sums, hipStream_t st) {
    return axis == 2 ? launch_last_bin_k<R, N, true>(spec, keep, P, axis, p, flags, sums, st)
                     : launch_last_bin_k<R, N, false>(spec, keep, P, axis, p, flags, sums, st);
}

#define PYL_PK_SIZES(X) \
    X(8) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160) X(192) X(256) \
    X(320) X(384) X(512) X(640) X(768) X(1024)

template <typename R>
int pk_last_pass_bin(const cplx<R>* spec, cplx<R>* keep, int N, int64_t P, int axis, int mas_p, int flags,
                     const PkSums& sums, hipStream_t st) {
    switch (N) {
#define PYL_CASE(n) \
    case n:         \
        return launch_last_bin<R, n>(spec, keep, P, axis, mas_p, flags, sums, st);
        PYL_PK_SIZES(PYL_CASE)
#undef PYL_CASE
        default:
            set_error("grid size %d not supported", N);
            return E_UNSUPPORTED;
    }
}
template int pk_last_pass_bin<float>(const cplx<float>*, cplx<float>*, int, int64_t, int, int, int, const PkSums&,
                                     hipStream_t);
template int pk_last_pass_bin<double>(const cplx<double>*, cplx<double>*, int, int64_t, int, int, int, const PkSums&,
                                      hipStream_t);

size_t last_bin_part_doubles(int N, int axis, int prec) {
    switch (N) {
#define PYL_CASE(n)                                                                                          \
    case n:                                                                                                  \
        if (prec == 8)                                                                                       \
            return axis == 2 ? (size_t)BinShape<double, n, true>().grid * BinShape<double, n, true>().RS     \
                             : (size_t)BinShape<double, n, false>().grid * BinShape<double, n, false>().RS;  \
        return axis == 2 ? (size_t)BinShape<float, n, true>().grid * BinShape<float, n, true>().RS           \
                         : (size_t)BinShape<float, n, false>().grid * BinShape<float, n, false>().RS;
        PYL_PK_SIZES(PYL_CASE)
#undef PYL_CASE
        default:
            return 0;
    }
}

template <typename R>
int last_bin_job(const cplx<R>* spec, int N, int64_t P, int axis, int mas_p, int flags, const PkSums& sums,
                 double* part, LastBinArgs<R>* a, ReduceArgs* r, unsigned* blocks) {
    switch (N) {
#define PYL_CASE(n)                                                                                      \
    case n:                                                                                              \
        return axis == 2 ? last_bin_job_k<R, n, true>(spec, nullptr, P, axis, mas_p, flags, sums, part, a, r, blocks) \
                         : last_bin_job_k<R, n, false>(spec, nullptr, P, axis, mas_p, flags, sums, part, a, r, blocks);
        PYL_PK_SIZES(PYL_CASE)
#undef PYL_CASE
        default:
            set_error("grid size %d not supported", N);
            return E_UNSUPPORTED;
    }
}

template <typename R>
int last_bin_batch(int N, int axis, const LastBinArgs<R>* tab, const unsigned* start, int nj, unsigned total,
                   const ReduceArgs* rtab, hipStream_t st) {
    switch (N) {
#define PYL_CASE(n)                                                                                      \
    case n:                                                                                              \
        return axis == 2 ? launch_last_bin_batch_k<R, n, true>(tab, start, nj, total, rtab, st)         \
                         : launch_last_bin_batch_k<R, n, false>(tab, start, nj, total, rtab, st);
        PYL_PK_SIZES(PYL_CASE)
#undef PYL_CASE
        default:
            set_error("grid size %d not supported", N);
            return E_UNSUPPORTED;
    }
}
template int last_bin_job<float>(const cplx<float>*, int, int64_t, int, int, int, const PkSums&, double*,
                                 LastBinArgs<float>*, ReduceArgs*, unsigned*);
template int last_bin_job<double>(const cplx<double>*, int, int64_t, int, int, int, const PkSums&, double*,
                                  LastBinArgs<double>*, ReduceArgs*, unsigned*);
template int last_bin_batch<float>(int, int, const LastBinArgs<float>*, const unsigned*, int, unsigned,
                                   const ReduceArgs*, hipStream_t);
template int last_bin_batch<double>(int, int, const LastBinArgs<double>*, const unsigned*, int, unsigned,
                                    const ReduceArgs*, hipStream_t);

// ------------------------------------------------------------------ finalize
size_t pk_sums_doubles(const PkDims& d) { return 5 * (size_t)d.n3() + 2 * (size_t)d.n1() + 2 * (size_t)d.n2() + 1; }

void pk_sums_carve(const PkDims& d, double* b, PkSums* s) {
    s->s3 = b;
    s->s1 = s->s3 + 4 * d.n3();
    s->s2 = s->s1 + d.n1();
    s->c3 = s->s2 + d.n2();
    s->c1 = s->c3 + d.n3();
    s->c2 = s->c1 + d.n1();
    s->status = (unsigned*)(s->c2 + d.n2());
}

// Nmodes of every histogram enabled in `flags` come from the binning kernel's counts and are
// compared, bin by bin, with the geometry (the 3-D comparison includes the DC bin, so it contains
// the total test of check_number_modes, Pk_library.pyx:90-102); histograms that were not binned
// report the geometry.

__device__ __forceinline__ void finalize_body(const FinArgs& A, Blk B) {
    // evaluate as written (no fused multiply-add), as the reference's Python arithmetic does:
    // the last 1-D bin needs kN^2 - k^2 == 0 exactly
#pragma clang fp contract(off)
    const int N = A.N, kmax = A.kmax, kmax_par = A.kmax_par, n2 = A.n2, flags = A.flags;
    const double BoxSize = A.BoxSize;
    const unsigned long long* __restrict__ n1c = A.n1c;
    const unsigned long long* __restrict__ n2c = A.n2c;
    const unsigned long long* __restrict__ n3c = A.n3c;
    const double* __restrict__ k3s = A.k3s;
    const double* __restrict__ s3 = A.s3;
    const double* __restrict__ s1 = A.s1;
    const double* __restrict__ s2 = A.s2;
    const double* __restrict__ c3 = A.c3;
    const double* __restrict__ c1 = A.c1;
    const double* __restrict__ c2 = A.c2;
    unsigned* __restrict__ status = A.status;
    double* __restrict__ out = A.out;
    const double kF = 2.0 * M_PI / BoxSize;
    const int middle = N / 2;
    const double kN = middle * kF;
    const double unit = (BoxSize / ((double)N * N)) * (BoxSize / ((double)N * N)) * (BoxSize / ((double)N * N));
    double* k3D = out;
    double* P0 = k3D + kmax;
    double* P2 = P0 + kmax;
    double* P4 = P2 + kmax;
    double* ph = P4 + kmax;
    double* N3 = ph + kmax;
    double* k1D = N3 + kmax;
    double* P1 = k1D + kmax_par;
    double* N1 = P1 + kmax_par;
    double* kpar = N1 + kmax_par;
    double* kper = kpar + n2;
    double* P2D = kper + n2;
    double* N2 = P2D + n2;
    const bool do1 = flags & PK_1D, do2 = flags & PK_2D;
    const int gi = B.x * blockDim.x + threadIdx.x, gs = B.gx * blockDim.x;
    unsigned bad = 0u;
    for (int i = gi; i <= kmax; i += gs) {
        const double n = c3[i];
        bad += n != (double)n3c[i] ? 1u : 0u;
        if (i == 0) continue;   // DC: counted and checked, not reported (Pk_library.pyx:412)
        const int o = i - 1;
        k3D[o] = (k3s[i] / n) * kF;
        P0[o] = (s3[4 * i + 0] / n) * unit;
        P2[o] = (s3[4 * i + 1] * 5.0 / n) * unit;
        P4[o] = (s3[4 * i + 2] * 9.0 / n) * unit;
        ph[o] = (s3[4 * i + 3] / n) * unit;
        N3[o] = n;
    }
    for (int i = gi; i <= kmax_par; i += gs) {
        const double n = do1 ? c1[i] : (double)n1c[i];
        if (do1) bad += n != (double)n1c[i] ? 1u : 0u;
        if (i == 0) continue;
        const int o = i - 1;
        const double k = ((double)i * n / n) * kF;
        const double kmaxper = sqrt(kN * kN - k * k);
        k1D[o] = k;
        P1[o] = s1[i] * unit * (M_PI * kmaxper * kmaxper / n) / ((2.0 * M_PI) * (2.0 * M_PI));
        N1[o] = n;
    }
    for (int i = gi; i < n2; i += gs) {
        const int kp = i % (kmax_par + 1), kq = i / (kmax_par + 1);
        const double n = do2 ? c2[i] : (double)n2c[i];
        if (do2) bad += n != (double)n2c[i] ? 1u : 0u;
        kpar[i] = 0.5 * (kp + kp + 1) * kF;
        kper[i] = 0.5 * (kq + kq + 1) * kF;
        P2D[i] = s2[i] * unit / n;
        N2[i] = n;
    }
    if (bad) atomicAdd(status, bad);
}
__global__ void k_pk_finalize(FinArgs a) { finalize_body(a, own_blk()); }
// batched: job blockIdx.y
__global__ void k_pk_finalize_b(const FinArgs* __restrict__ tab) {
    const FinArgs& a = tab[blockIdx.y];
    finalize_body(a, own_blk());
}

FinArgs finalize_args(const PkGeometry* g, const PkSums& sums, double BoxSize, int flags, double* out) {
    const PkDims& d = g->d;
    return FinArgs{d.N, d.kmax, d.kmax_par, d.n2(), BoxSize, flags, g->n1, g->n2, g->n3, g->k3sum, sums.s3, sums.s1,
                   sums.s2, sums.c3, sums.c1, sums.c2, sums.status, out};
}

static unsigned finalize_blocks(int N) {
    const PkDims d = pk_dims(N, half_len(N));
    return (unsigned)((std::max(d.kmax + 1, d.n2()) + 255) / 256);
}

int finalize_batch(int N, const FinArgs* tab, int nj, hipStream_t st) {
    hipLaunchKernelGGL(k_pk_finalize_b, dim3(finalize_blocks(N), (unsigned)nj), dim3(256), 0, st, tab);
    PYL_HIP(hipGetLastError());
    return OK;
}

int pk_finalize(const PkGeometry* g, const PkSums& sums, double BoxSize, int flags, double* out, hipStream_t st) {
    const FinArgs fa = finalize_args(g, sums, BoxSize, flags, out);
    hipLaunchKernelGGL(k_pk_finalize, dim3(finalize_blocks(g->d.N)), dim3(256), 0, st, fa);
    PYL_HIP(hipGetLastError());
    return OK;
}

}  // namespace pyl
