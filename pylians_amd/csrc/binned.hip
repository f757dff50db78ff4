// Cell-binned mass assignment for gfx950: no global atomics on the grid.
//
// 1. k_bin      particles -> fixed-capacity buckets, one per TILE^3 block of cells, keyed by the
//               tile holding the particle's lowest stencil cell (wave-aggregated bucket counters).
//               Particles beyond a bucket's capacity go to an overflow list.
// 2. k_ov_*     the overflow list is grouped by tile with an exact count + scan (empty in the
//               common case: the kernels then only read a zero counter).
// 3. k_tile     one workgroup per tile accumulates all stencil weights of its particles into an
//               LDS block of (TILE+S-1)^3 float64 cells (ds_add_f64: the per-cell sum is exact to
//               float64, the "compensated" accumulator), then writes the cells no other tile can
//               touch straight to the grid (one rounding) and hands the boundary cells (low S-1
//               layers and the S-1 spill layers beyond the high faces) to a float64 buffer.
// 4. k_merge    every tile sums the float64 partials of its low S-1 layers from itself and its up
//               to 7 lower neighbours and writes them (one rounding).
// Output modes: accumulate into the existing grid (MAS_library.MA semantics), overwrite, or the
// overdensity delta = rho * a - 1 (a = N^3 / mass) evaluated in float64 before the single rounding.
#include <math.h>

#include "deposit.hpp"
#include "profile.hpp"
#include "util.hpp"

namespace pyl {

constexpr int TILE = 16;

// Lowest stencil cell (unwrapped) and the float64 1-D weights; the same arithmetic as the
// stencils of deposit.hip (MAS_library.pyx:151-157, 290, 393-399, 486-492).
template <int S> struct Axis;
template <> struct Axis<1> {
    static __device__ __forceinline__ long long eval(double d, double* w) {
        w[0] = 1.0;
        return (long long)floor(d + 0.5);
    }
};
template <> struct Axis<2> {
    static __device__ __forceinline__ long long eval(double d, double* w) {
        const double b = floor(d);
        const double u = d - b;
        w[0] = 1.0 - u;
        w[1] = u;
        return (long long)b;
    }
};
template <> struct Axis<3> {
    static __device__ __forceinline__ long long eval(double d, double* w) {
        const double mn = floor(d - 1.5);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double s = fabs(mn + (j + 1) - d);
            w[j] = s < 0.5 ? 0.75 - s * s : (s < 1.5 ? 0.5 * (1.5 - s) * (1.5 - s) : 0.0);
        }
        return (long long)mn + 1;
    }
};
template <> struct Axis<4> {
    static __device__ __forceinline__ long long eval(double d, double* w) {
        const double mn = floor(d - 2.0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double s = fabs(mn + (j + 1) - d);
            w[j] = s < 1.0 ? (4.0 - 6.0 * s * s + 3.0 * s * s * s) / 6.0
                           : (s < 2.0 ? (2.0 - s) * (2.0 - s) * (2.0 - s) / 6.0 : 0.0);
        }
        return (long long)mn + 1;
    }
};

__device__ __forceinline__ int wrapN(long long b, int N) {
    long long r = b % N;
    return (int)(r < 0 ? r + N : r);
}

template <int S> struct Layout {
    static constexpr int M = TILE + S - 1;  // extended block edge
    static constexpr int H = S - 1;         // halo thickness
    static constexpr int K = TILE - H;      // interior edge
    static constexpr int XSZ = 2 * H * M * M;
    static constexpr int YSZ = K * 2 * H * M;
    static constexpr int ZSZ = K * K * 2 * H;
    static constexpr int BSZ = XSZ + YSZ + ZSZ;  // boundary cells per tile
    static constexpr int SHELL = TILE * TILE * TILE - K * K * K;
    // compact index of a boundary cell (i, j, k) of the extended block
    static __device__ __forceinline__ int bidx(int i, int j, int k) {
        if (i < H || i >= TILE) {
            const int ii = i < H ? i : i - TILE + H;
            return (ii * M + j) * M + k;
        }
        const int i2 = i - H;
        if (j < H || j >= TILE) {
            const int jj = j < H ? j : j - TILE + H;
            return XSZ + (i2 * 2 * H + jj) * M + k;
        }
        const int j2 = j - H;
        const int kk = k < H ? k : k - TILE + H;
        return XSZ + YSZ + (i2 * K + j2) * 2 * H + kk;
    }
    static __device__ __forceinline__ void bcell(int idx, int& i, int& j, int& k) {
        if (idx < XSZ) {
            const int ii = idx / (M * M), r = idx % (M * M);
            i = ii < H ? ii : ii - H + TILE;
            j = r / M;
            k = r % M;
        } else if (idx < XSZ + YSZ) {
            const int r = idx - XSZ, i2 = r / (2 * H * M), q = r % (2 * H * M), jj = q / M;
            i = i2 + H;
            j = jj < H ? jj : jj - H + TILE;
            k = q % M;
        } else {
            const int r = idx - XSZ - YSZ, i2 = r / (K * 2 * H), q = r % (K * 2 * H), j2 = q / (2 * H),
                      kk = q % (2 * H);
            i = i2 + H;
            j = j2 + H;
            k = kk < H ? kk : kk - H + TILE;
        }
    }
};

struct BinArgs {
    const float* pos;
    const float* W;
    long long n;
    int N, nt;           // grid edge, tiles per edge
    double inv;          // N / BoxSize
    unsigned* fill;      // [ntiles]
    long long C;         // bucket capacity (records)
    float* bucket;       // [ntiles * C * R]
    unsigned* ov_count;  // [1]
    long long* ov_idx;   // [n]
    unsigned* ov_tile;   // [n]
};

template <int S, bool HASW>
__global__ __launch_bounds__(256) void k_bin(BinArgs a) {
    constexpr int R = HASW ? 4 : 3;
    const int lane = threadIdx.x & 63;
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long first = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    // every lane of a wave runs the same number of iterations (ballots below need them all)
    const long long iters = (a.n + stride - 1) / stride;
    for (long long it = 0; it < iters; ++it) {
        const long long p = first + it * stride;
        const bool valid = p < a.n;
        float x = 0.f, y = 0.f, z = 0.f, w = 0.f;
        int t = -1;
        if (valid) {
            x = a.pos[3 * p];
            y = a.pos[3 * p + 1];
            z = a.pos[3 * p + 2];
            if (HASW) w = a.W[p];
            double ww[4];
            const int bx = wrapN(Axis<S>::eval((double)x * a.inv, ww), a.N);
            const int by = wrapN(Axis<S>::eval((double)y * a.inv, ww), a.N);
            const int bz = wrapN(Axis<S>::eval((double)z * a.inv, ww), a.N);
            t = ((bx / TILE) * a.nt + by / TILE) * a.nt + bz / TILE;
        }
        unsigned long long active = __ballot(valid);
        long long slot = -1;
        while (active) {
            const int leader = __ffsll((long long)active) - 1;
            const int key = __shfl(t, leader, 64);
            const unsigned long long m = __ballot(t == key) & active;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(&a.fill[key], (unsigned)__popcll(m));
            base = __shfl(base, leader, 64);
            if ((m >> lane) & 1ull) slot = (long long)base + __popcll(m & ((1ull << lane) - 1ull));
            active &= ~m;
        }
        const bool inb = valid && slot < a.C;
        if (inb) {
            float* r = a.bucket + ((long long)t * a.C + slot) * R;
            r[0] = x;
            r[1] = y;
            r[2] = z;
            if (HASW) r[3] = w;
        }
        const unsigned long long ovm = __ballot(valid && !inb);
        if (ovm) {
            const int leader = __ffsll((long long)ovm) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(a.ov_count, (unsigned)__popcll(ovm));
            base = __shfl(base, leader, 64);
            if ((ovm >> lane) & 1ull) {
                const unsigned o = base + __popcll(ovm & ((1ull << lane) - 1ull));
                a.ov_idx[o] = p;
                a.ov_tile[o] = (unsigned)t;
            }
        }
    }
}

// overflow grouping: count per tile, exclusive scan, scatter (records copied from the inputs)
__global__ __launch_bounds__(256) void k_ov_count(const unsigned* __restrict__ ov_count,
                                                   const unsigned* __restrict__ ov_tile, unsigned* __restrict__ ocnt) {
    const long long n = *ov_count;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        atomicAdd(&ocnt[ov_tile[i]], 1u);
}

__global__ __launch_bounds__(1024) void k_scan(const unsigned* __restrict__ cnt, unsigned* __restrict__ off, int n) {
    __shared__ unsigned part[1024];
    const int tid = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int lo = tid * per, hi = min(n, lo + per);
    unsigned s = 0;
    for (int i = lo; i < hi; ++i) s += cnt[i];
    part[tid] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        unsigned v = tid >= d ? part[tid - d] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    unsigned run = tid ? part[tid - 1] : 0u;
    for (int i = lo; i < hi; ++i) {
        off[i] = run;
        run += cnt[i];
    }
    if (tid == 1023) off[n] = part[1023];
}

template <bool HASW>
__global__ __launch_bounds__(256) void k_ov_scatter(const unsigned* __restrict__ ov_count,
                                                     const long long* __restrict__ ov_idx,
                                                     const unsigned* __restrict__ ov_tile,
                                                     const unsigned* __restrict__ ooff, unsigned* __restrict__ ocur,
                                                     const float* __restrict__ pos, const float* __restrict__ W,
                                                     float* __restrict__ grouped) {
    constexpr int R = HASW ? 4 : 3;
    const long long n = *ov_count;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const unsigned t = ov_tile[i];
        const long long p = ov_idx[i];
        const unsigned s = ooff[t] + atomicAdd(&ocur[t], 1u);
        float* r = grouped + (long long)s * R;
        r[0] = pos[3 * p];
        r[1] = pos[3 * p + 1];
        r[2] = pos[3 * p + 2];
        if (HASW) r[3] = W[p];
    }
}

enum { OUT_ACC = 0, OUT_SET = 1, OUT_DELTA = 2 };

template <typename G, int MODE>
__device__ __forceinline__ void put(G* __restrict__ grid, long long g, double rho, double a) {
    if (MODE == OUT_ACC)
        grid[g] = (G)((double)grid[g] + rho);
    else if (MODE == OUT_SET)
        grid[g] = (G)rho;
    else
        grid[g] = (G)(rho * a - 1.0);
}

struct TileArgs {
    const float* bucket;
    const unsigned* fill;
    long long C;
    const float* grouped;
    const unsigned* ooff;
    const unsigned* ocnt;
    int N, nt;
    double inv;
    double mult;          // weight multiplier
    void* grid;
    double* bnd;          // [ntiles * BSZ]
    const double* dmass;  // device mass (nullable) for OUT_DELTA
    double mass;
};

template <int S, bool HASW>
__device__ __forceinline__ void accumulate(double* acc, const float* r, int N, int tx, int ty, int tz, double inv,
                                           double mult) {
    constexpr int M = Layout<S>::M;
    double w0[S], w1[S], w2[S];
    const int l0 = wrapN(Axis<S>::eval((double)r[0] * inv, w0), N) - tx * TILE;
    const int l1 = wrapN(Axis<S>::eval((double)r[1] * inv, w1), N) - ty * TILE;
    const int l2 = wrapN(Axis<S>::eval((double)r[2] * inv, w2), N) - tz * TILE;
    const double ww = HASW ? (double)r[3] * mult : mult;
#pragma unroll
    for (int l = 0; l < S; ++l)
#pragma unroll
        for (int m = 0; m < S; ++m) {
            const double wlm = w0[l] * w1[m] * ww;
            double* row = acc + ((l0 + l) * M + (l1 + m)) * M + l2;
#pragma unroll
            for (int q = 0; q < S; ++q) atomicAdd(row + q, wlm * w2[q]);
        }
}

template <int S, typename G, bool HASW, int MODE>
__global__ __launch_bounds__(256) void k_tile(TileArgs a) {
    using Lt = Layout<S>;
    constexpr int M = Lt::M, H = Lt::H, K = Lt::K;
    constexpr int R = HASW ? 4 : 3;
    __shared__ double acc[M * M * M];
    const int t = blockIdx.x;
    const int tx = t / (a.nt * a.nt), ty = (t / a.nt) % a.nt, tz = t % a.nt;
    for (int i = threadIdx.x; i < M * M * M; i += 256) acc[i] = 0.0;
    __syncthreads();
    const long long cnt = min((long long)a.fill[t], a.C);
    const float* rec = a.bucket + (long long)t * a.C * R;
    for (long long p = threadIdx.x; p < cnt; p += 256)
        accumulate<S, HASW>(acc, rec + p * R, a.N, tx, ty, tz, a.inv, a.mult);
    const long long o0 = a.ooff[t], oc = a.ocnt[t];
    for (long long p = threadIdx.x; p < oc; p += 256)
        accumulate<S, HASW>(acc, a.grouped + (o0 + p) * R, a.N, tx, ty, tz, a.inv, a.mult);
    __syncthreads();
    G* grid = (G*)a.grid;
    const double am = (MODE == OUT_DELTA) ? (double)a.N * a.N * a.N / (a.dmass ? a.dmass[0] : a.mass) : 0.0;
    const long long N = a.N;
    // cells only this tile touches: local [H, TILE)^3
    for (int idx = threadIdx.x; idx < K * K * K; idx += 256) {
        const int i = H + idx / (K * K), j = H + (idx / K) % K, k = H + idx % K;
        const long long g = (((long long)tx * TILE + i) * N + ty * TILE + j) * N + tz * TILE + k;
        put<G, MODE>(grid, g, acc[(i * M + j) * M + k], am);
    }
    if (H > 0) {
        double* b = a.bnd + (long long)t * Lt::BSZ;
        for (int idx = threadIdx.x; idx < Lt::BSZ; idx += 256) {
            int i, j, k;
            Lt::bcell(idx, i, j, k);
            b[idx] = acc[(i * M + j) * M + k];
        }
    }
}

template <int S, typename G, int MODE>
__global__ __launch_bounds__(256) void k_merge(TileArgs a) {
    using Lt = Layout<S>;
    constexpr int H = Lt::H, K = Lt::K;
    const int t = blockIdx.x;
    const int nt = a.nt;
    const int tx = t / (nt * nt), ty = (t / nt) % nt, tz = t % nt;
    G* grid = (G*)a.grid;
    const double am = (MODE == OUT_DELTA) ? (double)a.N * a.N * a.N / (a.dmass ? a.dmass[0] : a.mass) : 0.0;
    const long long N = a.N;
    constexpr int XL = H * TILE * TILE, YL = K * H * TILE;
    for (int idx = threadIdx.x; idx < Lt::SHELL; idx += 256) {
        int i, j, k;
        if (idx < XL) {
            i = idx / (TILE * TILE);
            j = (idx / TILE) % TILE;
            k = idx % TILE;
        } else if (idx < XL + YL) {
            const int r = idx - XL;
            i = H + r / (H * TILE);
            j = (r / TILE) % H;
            k = r % TILE;
        } else {
            const int r = idx - XL - YL;
            i = H + r / (K * H);
            j = H + (r / H) % K;
            k = r % H;
        }
        double s = 0.0;
        for (int dx = 0; dx <= (i < H ? 1 : 0); ++dx)
            for (int dy = 0; dy <= (j < H ? 1 : 0); ++dy)
                for (int dz = 0; dz <= (k < H ? 1 : 0); ++dz) {
                    const int sx = (tx - dx + nt) % nt, sy = (ty - dy + nt) % nt, sz = (tz - dz + nt) % nt;
                    const long long st = ((long long)sx * nt + sy) * nt + sz;
                    s += a.bnd[st * Lt::BSZ + Lt::bidx(i + dx * TILE, j + dy * TILE, k + dz * TILE)];
                }
        const long long g = (((long long)tx * TILE + i) * N + ty * TILE + j) * N + tz * TILE + k;
        put<G, MODE>(grid, g, s, am);
    }
}

// ------------------------------------------------------------------ host side
bool binned_supported(int axes, int dims) { return axes == 3 && dims % TILE == 0 && dims >= 2 * TILE; }

template <int S, typename G, bool HASW, int MODE>
static int run_binned(const float* pos, const float* W, int64_t n, int dims, double BoxSize, double mult, void* grid,
                      const double* dmass, double mass, hipStream_t st) {
    using Lt = Layout<S>;
    constexpr int R = HASW ? 4 : 3;
    int rc = OK;
    const int nt = dims / TILE;
    const long long ntiles = (long long)nt * nt * nt;
    const long long avg = (n + ntiles - 1) / ntiles;
    const long long C = ((2 * avg + 64) + 15) / 16 * 16;
    // counters: fill | ocnt | ocur | ooff (ntiles+1) | ov_count
    const size_t ncnt = 4 * (size_t)ntiles + 2;
    unsigned* cnt = (unsigned*)workspace(WS_BINS, sizeof(unsigned) * ncnt + 64, &rc);
    if (rc) return rc;
    unsigned *fill = cnt, *ocnt = cnt + ntiles, *ocur = cnt + 2 * ntiles, *ooff = cnt + 3 * ntiles,
             *ov_count = cnt + 4 * ntiles + 1;
    float* bucket = (float*)workspace(WS_AUX, sizeof(float) * (size_t)ntiles * C * R, &rc);
    if (rc) return rc;
    char* ov = (char*)workspace(WS_AUX2, (size_t)n * (8 + 4 + 4 * R) + 256, &rc);
    if (rc) return rc;
    long long* ov_idx = (long long*)ov;
    unsigned* ov_tile = (unsigned*)(ov + 8 * (size_t)n);
    float* grouped = (float*)(ov + 12 * (size_t)n + 64);
    double* bnd = nullptr;
    if (Lt::BSZ > 0) {
        bnd = (double*)workspace(WS_AUX3, sizeof(double) * (size_t)ntiles * Lt::BSZ, &rc);
        if (rc) return rc;
    }
    rc = zero_fill(cnt, sizeof(unsigned) * ncnt, st);
    if (rc) return rc;
    {
        StageTimer tm("deposit_bin", st);
        BinArgs ba{pos, W, (long long)n, dims, nt, dims / BoxSize, fill, C, bucket, ov_count, ov_idx, ov_tile};
        long long blocks = (n + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        if (blocks < 1) blocks = 1;
        hipLaunchKernelGGL((k_bin<S, HASW>), dim3((unsigned)blocks), dim3(256), 0, st, ba);
        PYL_HIP(hipGetLastError());
    }
    {
        StageTimer tm("deposit_overflow", st);
        hipLaunchKernelGGL(k_ov_count, dim3(256), dim3(256), 0, st, ov_count, ov_tile, ocnt);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, ocnt, ooff, (int)ntiles);
        hipLaunchKernelGGL((k_ov_scatter<HASW>), dim3(256), dim3(256), 0, st, ov_count, ov_idx, ov_tile, ooff, ocur,
                           pos, W, grouped);
        PYL_HIP(hipGetLastError());
    }
    TileArgs ta{bucket, fill, C, grouped, ooff, ocnt, dims, nt, dims / BoxSize, mult, grid, bnd, dmass, mass};
    {
        StageTimer tm("deposit_tiles", st);
        hipLaunchKernelGGL((k_tile<S, G, HASW, MODE>), dim3((unsigned)ntiles), dim3(256), 0, st, ta);
        PYL_HIP(hipGetLastError());
    }
    if (Lt::BSZ > 0) {
        StageTimer tm("deposit_merge", st);
        hipLaunchKernelGGL((k_merge<S, G, MODE>), dim3((unsigned)ntiles), dim3(256), 0, st, ta);
        PYL_HIP(hipGetLastError());
    }
    return OK;
}

template <int S, typename G, int MODE>
static int dispatch_w(const float* pos, const float* W, int64_t n, int dims, double BoxSize, double mult, void* grid,
                      const double* dmass, double mass, hipStream_t st) {
    return W ? run_binned<S, G, true, MODE>(pos, W, n, dims, BoxSize, mult, grid, dmass, mass, st)
             : run_binned<S, G, false, MODE>(pos, W, n, dims, BoxSize, mult, grid, dmass, mass, st);
}

template <typename G, int MODE>
static int dispatch_s(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, double mult,
                      void* grid, const double* dmass, double mass, hipStream_t st) {
    switch (scheme) {
        case 1: return dispatch_w<1, G, MODE>(pos, W, n, dims, BoxSize, mult, grid, dmass, mass, st);
        case 2: return dispatch_w<2, G, MODE>(pos, W, n, dims, BoxSize, mult, grid, dmass, mass, st);
        case 3: return dispatch_w<3, G, MODE>(pos, W, n, dims, BoxSize, mult, grid, dmass, mass, st);
        case 4: return dispatch_w<4, G, MODE>(pos, W, n, dims, BoxSize, mult, grid, dmass, mass, st);
        default: set_error("unknown scheme %d", scheme); return E_ARG;
    }
}

// mode: 0 accumulate, 1 overwrite, 2 overdensity (mass = dmass[0] if dmass else mass)
int deposit_binned(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, bool f64,
                   int mode, double mult, void* grid, const double* dmass, double mass, hipStream_t st) {
    PYL_CHECK(binned_supported(3, dims), E_UNSUPPORTED, "binned deposit needs 3-D grids with dims %% %d == 0", TILE);
    PYL_CHECK(n > 0, E_ARG, "binned deposit needs particles");
    PYL_CHECK(n < (1LL << 31), E_UNSUPPORTED, "binned deposit: at most 2^31 particles per call");
    if (f64) {
        if (mode == OUT_ACC) return dispatch_s<double, OUT_ACC>(pos, W, n, dims, BoxSize, scheme, mult, grid, dmass, mass, st);
        if (mode == OUT_SET) return dispatch_s<double, OUT_SET>(pos, W, n, dims, BoxSize, scheme, mult, grid, dmass, mass, st);
        return dispatch_s<double, OUT_DELTA>(pos, W, n, dims, BoxSize, scheme, mult, grid, dmass, mass, st);
    }
    if (mode == OUT_ACC) return dispatch_s<float, OUT_ACC>(pos, W, n, dims, BoxSize, scheme, mult, grid, dmass, mass, st);
    if (mode == OUT_SET) return dispatch_s<float, OUT_SET>(pos, W, n, dims, BoxSize, scheme, mult, grid, dmass, mass, st);
    return dispatch_s<float, OUT_DELTA>(pos, W, n, dims, BoxSize, scheme, mult, grid, dmass, mass, st);
}

}  // namespace pyl
