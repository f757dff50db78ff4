// Cell-binned mass assignment for gfx950: no global atomics on the grid, every cell written once.
//
// The grid is cut into TILE^3 tiles.  A particle belongs to the tile of its lowest stencil cell; its
// halo class is the 3-bit mask of the axes along which its stencil reaches into the next tile (local
// lowest cell > TILE - S).  Buckets are segmented by (tile, class).
// 1. k_bin    particles -> bucket segments.  A workgroup takes BT * KP consecutive particles at a time,
//             counts them per (tile, class) key in an LDS hash table (the positions of the next
//             step are already loading), reserves room with one global atomic per distinct key and
//             writes the records (position, and weight).  Records beyond
//             a segment's capacity go to an overflow list.
// 2. k_ov_*   the overflow list is grouped by tile with an exact count + scan (empty in the common
//             case: the kernels then only read a zero counter).
// 3. k_tile   one workgroup per tile accumulates into TILE^3 float64 LDS cells (ds_add_f64: the per-
//             cell sum is exact to float64, the "compensated" accumulator) the stencil weights that
//             fall inside the tile, from its own particles and from those of its 7 lower neighbours
//             whose class reaches it, then writes every cell once with one rounding.
// Output modes: accumulate into the existing grid (MAS_library.MA semantics), overwrite, or the
// overdensity delta = rho * a - 1 (a = N^3 / mass) evaluated in float64 before the single rounding.
#include <math.h>

#include <algorithm>

#include "binned.hpp"
#include "deposit.hpp"
#include "profile.hpp"
#include "util.hpp"

namespace pyl {

constexpr int BT = 512;    // k_bin: threads per workgroup
constexpr int KP = 4;      // k_bin: particles per thread per step
constexpr int HT = 4096;   // k_bin: LDS hash slots (> BT * KP, so an insertion always finds a slot)
constexpr int TT = 512;    // k_tile: threads per workgroup
// particles per cell from which CIC tiles accumulate in register runs (measured at 512^3 particles: 8 per
// cell 0.94 -> 0.78 ms, 4.1: 0.82 -> 0.75, 2.4: equal, 1.5: 0.75 -> 0.81; NGP gains nothing)
constexpr double RUN_PPC = 3.0;
// row stride of the float64 LDS accumulator of a tile: TILE + 1 doubles, so that particles one row
// apart in y (neighbours in the input along z fill a row, then the next) land on LDS banks 34 apart
// (measured on the 512^3 lattice: k_tile 0.88 -> 0.77 ms; strides 18/20/24 and padded planes were slower)
constexpr int TROW = TILE + 1;
static_assert(HT % BT == 0, "hash slots per thread");
constexpr int HBITS = 12;   // log2(HT): the slot of a key is the top HBITS bits of its multiplicative hash
static_assert(HT == 1 << HBITS, "HT must be 2^HBITS");
static_assert(HT > BT * KP, "every key of a step must find a hash slot");

// Lowest stencil cell (unwrapped) and the float64 1-D weights; the same arithmetic as the
// stencils of deposit.hip (MAS_library.pyx:151-157, 290, 393-399, 486-492).
template <int S> struct Axis;
template <> struct Axis<1> {
    static __device__ __forceinline__ double eval(double d, double* w) {
        w[0] = 1.0;
        return floor(d + 0.5);
    }
};
template <> struct Axis<2> {
    static __device__ __forceinline__ double eval(double d, double* w) {
        const double b = floor(d);
        const double u = d - b;
        w[0] = 1.0 - u;
        w[1] = u;
        return b;
    }
};
template <> struct Axis<3> {
    static __device__ __forceinline__ double eval(double d, double* w) {
        const double mn = floor(d - 1.5);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double s = fabs(mn + (j + 1) - d);
            w[j] = s < 0.5 ? 0.75 - s * s : (s < 1.5 ? 0.5 * (1.5 - s) * (1.5 - s) : 0.0);
        }
        return mn + 1.0;
    }
};
template <> struct Axis<4> {
    static __device__ __forceinline__ double eval(double d, double* w) {
        const double mn = floor(d - 2.0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double s = fabs(mn + (j + 1) - d);
            w[j] = s < 1.0 ? (4.0 - 6.0 * s * s + 3.0 * s * s * s) / 6.0
                           : (s < 2.0 ? (2.0 - s) * (2.0 - s) * (2.0 - s) / 6.0 : 0.0);
        }
        return mn + 1.0;
    }
};

// The same stencils with the 1-D weights evaluated in float32 from the float64 cell fraction
// (CIC: u and 1 - u in float64, each rounded once; TSC/PCS: the distance |cell - d| and its
// distance to the stencil's outer edge in float64, each rounded once, then the polynomials in
// float32, so that no weight loses precision to cancellation); the lowest cell is the float64 one.
template <int S> struct Axis32;
template <> struct Axis32<1> {
    static __device__ __forceinline__ double eval(double d, float* w) {
        w[0] = 1.0f;
        return floor(d + 0.5);
    }
};
template <> struct Axis32<2> {
    static __device__ __forceinline__ double eval(double d, float* w) {
        const double b = floor(d);
        const double u = d - b;
        w[0] = (float)(1.0 - u);
        w[1] = (float)u;
        return b;
    }
};
template <> struct Axis32<3> {
    static __device__ __forceinline__ double eval(double d, float* w) {
        const double mn = floor(d - 1.5);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double s = fabs(mn + (j + 1) - d);
            const float sf = (float)s, t = (float)(1.5 - s);   // distance to the outer edge, exact first
            w[j] = s < 0.5 ? 0.75f - sf * sf : (s < 1.5 ? 0.5f * t * t : 0.0f);
        }
        return mn + 1.0;
    }
};
template <> struct Axis32<4> {
    static __device__ __forceinline__ double eval(double d, float* w) {
        const double mn = floor(d - 2.0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double s = fabs(mn + (j + 1) - d);
            const float sf = (float)s, t = (float)(2.0 - s);
            w[j] = s < 1.0 ? (4.0f - 6.0f * sf * sf + 3.0f * sf * sf * sf) * (1.0f / 6.0f)
                           : (s < 2.0 ? t * t * t * (1.0f / 6.0f) : 0.0f);
        }
        return mn + 1.0;
    }
};

// periodic cell index of an integral float64 cell coordinate b (positions inside the box give
// 0 <= b < N directly, a stencil edge at most one box away; others reduce with a float64 floor)
__device__ __forceinline__ int wrapN(double b, int N) {
    // v_cvt_i32_f64 clamps out-of-range values to INT_MIN / INT_MAX and turns NaN into 0, so one
    // integer test takes every cell inside the grid (and NaN to cell 0, as below)
    int i;
    asm("v_cvt_i32_f64 %0, %1" : "=v"(i) : "v"(b));
    if ((unsigned)i < (unsigned)N) return i;
    const double n = (double)N;
    if (b < 0.0 && b >= -n) return (int)(b + n);
    if (b >= n && b < 2.0 * n) return (int)(b - n);
    int r = (int)(b - n * floor(b / n));
    if (r >= N) r -= N;
    if (r < 0) r += N;
    return (unsigned)r < (unsigned)N ? r : 0;   // non-finite input: stay inside the grid
}


template <int S>
__device__ __forceinline__ int halo_class(int bx, int by, int bz) {
    return ((bx % TILE) > TILE - S ? 1 : 0) | ((by % TILE) > TILE - S ? 2 : 0) | ((bz % TILE) > TILE - S ? 4 : 0);
}

// AGG: 0 every lane counts itself, 1 runs of equal keys are counted once, 2 a.agg decides (batched
// launches, whose jobs differ)
template <int S, bool HASW, int AGG>
__device__ __forceinline__ void bin_body(const BinArgs& a, Blk B) {
    constexpr int R = HASW ? 4 : 3;
    constexpr long long PS = (long long)BT * KP;   // particles per step
    // a slot holds a key of the step, then (after the reservations) the first index of the key's records
    __shared__ int hkey[HT];
    __shared__ unsigned hcnt[HT];
    __shared__ int hused[BT * KP];   // slots that hold a key this step, in insertion order
    __shared__ unsigned nused;
    __shared__ int scap[8], scofs[8];
    const int tid = threadIdx.x, lane = tid & 63;
    if (tid < 8) {
        scap[tid] = a.cap[tid];
        scofs[tid] = a.cofs[tid];
    }
    for (int i = tid; i < HT; i += BT) {
        hkey[i] = -1;
        hcnt[i] = 0u;
    }
    if (tid == 0) nused = 0u;
    __syncthreads();
    const long long nstep = (a.n + PS - 1) / PS;
    // particle k of this thread in a step: step * PS + k * BT + tid (coalesced for each k)
    float px[KP], py[KP], pz[KP], pw[KP];
    auto load = [&](long long step) {
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const long long p = step * PS + k * BT + tid;
            if (p < a.n) {
                float c[3] = {a.pos[3 * p], a.pos[3 * p + 1], a.pos[3 * p + 2]};
                if (a.rsd.vel) {
                    const int ax = a.rsd.axis;
                    c[ax] = rsd_shift(c[ax], a.rsd.vel[3 * p + ax], a.rsd.fac, a.rsd.box);
                }
                px[k] = c[0];
                py[k] = c[1];
                pz[k] = c[2];
                pw[k] = HASW ? a.W[p] : 0.f;
            }
        }
    };
    // each XCD (workgroups go round-robin over the XCDS of the chip) takes one contiguous range of
    // steps, so that records of one segment written at about the same time share an L2, and each
    // workgroup of the XCD a contiguous part of it: consecutive steps of one workgroup append to the
    // same segments, and workgroups running at the same time reserve in different segments (measured
    // 3 % faster than interleaving the workgroups step by step)
    const long long GX = B.gx / XCDS, per = (nstep + XCDS - 1) / XCDS;
    const long long s_beg = (long long)(B.x % XCDS) * per;
    const long long s_end = min(nstep, s_beg + per);
    const long long bper = (per + GX - 1) / GX;
    const long long b_beg = s_beg + (long long)(B.x / XCDS) * bper;
    const long long b_end = min(s_end, b_beg + bper);
    long long step = b_beg;
    if (step < b_end) load(step);
    for (; step < b_end; ++step) {
        float x[KP], y[KP], z[KP], w[KP];
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            x[k] = px[k];
            y[k] = py[k];
            z[k] = pz[k];
            w[k] = pw[k];
        }
        if (step + 1 < b_end) load(step + 1);   // in flight during this step
        int key[KP], slot[KP];
        unsigned rank[KP];
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            const long long p = step * PS + k * BT + tid;
            const bool valid = p < a.n;
            key[k] = -1;
            slot[k] = 0;
            rank[k] = 0u;
            if (valid) {
                double ww[4];
                const int bx = wrapN(Axis<S>::eval((double)x[k] * a.inv, ww), a.N);
                const int by = wrapN(Axis<S>::eval((double)y[k] * a.inv, ww), a.N);
                const int bz = wrapN(Axis<S>::eval((double)z[k] * a.inv, ww), a.N);
                key[k] = (((bx / TILE) * a.nt + by / TILE) * a.nt + bz / TILE) * 8 + halo_class<S>(bx, by, bz);
            }
            // at most BT * KP distinct keys per step in HT slots: the probe always ends
            auto probe = [&](int kk) {
                int s0 = (int)(((unsigned)kk * 2654435761u) >> (32 - HBITS));
                for (;;) {
                    const int prev = atomicCAS(&hkey[s0], -1, kk);
                    if (prev == -1) hused[atomicAdd(&nused, 1u)] = s0;
                    if (prev == -1 || prev == kk) return s0;
                    s0 = (s0 + 1) & (HT - 1);
                }
            };
            if (AGG == 1 || (AGG == 2 && a.agg)) {
                // (uniform over the launch: every lane runs the shuffles) runs of equal keys in
                // consecutive lanes: the first lane of a run probes the table and counts the whole run,
                // the others take the slot and their rank from it.  Invalid lanes (key -1) form runs of
                // their own, which are skipped below.
                const int lane = tid & 63;
                const int kprev = __shfl_up(key[k], 1);
                const unsigned long long hm = __ballot(lane == 0 || kprev != key[k]);
                const unsigned long long le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);   // lanes <= lane
                const int start = 63 - __clzll((long long)(hm & le));   // lane 0 always starts a run
                const unsigned long long gt = hm & ~le;
                const int end = gt ? __ffsll((long long)gt) - 1 : 64;
                int s0 = 0;
                unsigned base = 0u;
                if (valid && start == lane) {
                    s0 = probe(key[k]);
                    base = atomicAdd(&hcnt[s0], (unsigned)(end - start));
                }
                slot[k] = __shfl(s0, start);
                rank[k] = __shfl(base, start) + (unsigned)(lane - start);
            } else if (valid) {
                slot[k] = probe(key[k]);
                rank[k] = atomicAdd(&hcnt[slot[k]], 1u);
            }
        }
        __syncthreads();
        // one reservation per key of the step: its records go to [base, base + hcnt) of its segment;
        // the base replaces the key in the slot
        const unsigned nu = nused;
        for (unsigned i = tid; i < nu; i += BT) {
            const int sl = hused[i];
            hkey[sl] = (int)atomicAdd(&a.fill[hkey[sl]], hcnt[sl]);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KP; ++k) {
            if (key[k] < 0) continue;
            const long long p = step * PS + k * BT + tid;
            const int c = key[k] & 7;
            const unsigned idx = (unsigned)hkey[slot[k]] + rank[k];
            if (idx < (unsigned)scap[c]) {
                float* r = a.bucket + ((long long)(key[k] >> 3) * a.CT + scofs[c] + idx) * R;
                r[0] = x[k];
                r[1] = y[k];
                r[2] = z[k];
                if (HASW) r[3] = w[k];
            } else {
                const unsigned o = atomicAdd(a.ov_count, 1u);
                a.ov_idx[o] = p;
                a.ov_key[o] = (unsigned)key[k];
            }
        }
        __syncthreads();
        for (unsigned i = tid; i < nu; i += BT) {
            const int sl = hused[i];
            hkey[sl] = -1;
            hcnt[sl] = 0u;
        }
        if (tid == 0) nused = 0u;
        __syncthreads();
    }
}

template <int S, bool HASW, int AGG>
__global__ __launch_bounds__(BT) void k_bin(BinArgs a) { bin_body<S, HASW, AGG>(a, own_blk()); }
template <int S, bool HASW>
__global__ __launch_bounds__(BT) void k_bin_b(const BinArgs* __restrict__ tab,
                                                 const unsigned* __restrict__ start, int nj) {
    PYL_BATCH_BLK(start, nj);
    bin_body<S, HASW, 2>(tab[bj_], blk_);
}

// overflow grouping: count per tile, exclusive scan, scatter (records copied from the inputs)

__device__ __forceinline__ void ov_count_body(const OvArgs& a, Blk B) {
    const long long n = *a.ov_count;
    for (long long i = (long long)B.x * blockDim.x + threadIdx.x; i < n; i += (long long)B.gx * blockDim.x)
        atomicAdd(&a.ocnt[a.ov_key[i] >> 3], 1u);
}
__global__ __launch_bounds__(256) void k_ov_count(OvArgs a) { ov_count_body(a, own_blk()); }
// batched: job blockIdx.y
__global__ __launch_bounds__(256) void k_ov_count_b(const OvArgs* __restrict__ tab) {
    const OvArgs& a = tab[blockIdx.y];
    ov_count_body(a, own_blk());
}

// (nothing overflowed: every group is empty and k_tile never reads an offset, so skip the scan)
__device__ __forceinline__ void scan_body(const unsigned* __restrict__ ov_count, const unsigned* __restrict__ cnt,
                                          unsigned* __restrict__ off, int n) {
    __shared__ unsigned part[1024];
    if (*ov_count == 0u) return;
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
__global__ __launch_bounds__(1024) void k_scan(OvArgs a) { scan_body(a.ov_count, a.ocnt, a.ooff, a.ntiles); }
// batched: one workgroup per job
__global__ __launch_bounds__(1024) void k_scan_b(const OvArgs* __restrict__ tab) {
    const OvArgs& a = tab[blockIdx.x];
    scan_body(a.ov_count, a.ocnt, a.ooff, a.ntiles);
}

template <bool HASW>
__device__ __forceinline__ void ov_scatter_body(const OvArgs& a, Blk B) {
    constexpr int R = HASW ? 4 : 3;
    const long long n = *a.ov_count;
    for (long long i = (long long)B.x * blockDim.x + threadIdx.x; i < n; i += (long long)B.gx * blockDim.x) {
        const unsigned key = a.ov_key[i], t = key >> 3;
        const long long p = a.ov_idx[i];
        const unsigned s = a.ooff[t] + atomicAdd(&a.ocur[t], 1u);
        float* r = a.grouped + (long long)s * R;
        float c[3] = {a.pos[3 * p], a.pos[3 * p + 1], a.pos[3 * p + 2]};
        if (a.rsd.vel) c[a.rsd.axis] = rsd_shift(c[a.rsd.axis], a.rsd.vel[3 * p + a.rsd.axis], a.rsd.fac, a.rsd.box);
        r[0] = c[0];
        r[1] = c[1];
        r[2] = c[2];
        if (HASW) r[3] = a.W[p];
        a.gcls[s] = key & 7u;
    }
}
template <bool HASW>
__global__ __launch_bounds__(256) void k_ov_scatter(OvArgs a) { ov_scatter_body<HASW>(a, own_blk()); }
template <bool HASW>
__global__ __launch_bounds__(256) void k_ov_scatter_b(const OvArgs* __restrict__ tab) {
    const OvArgs& a = tab[blockIdx.y];
    ov_scatter_body<HASW>(a, own_blk());
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


// the stencil weights of one particle that fall inside the tile with lowest cell (ox, oy, oz)
// (CLIP = false: a particle of the tile itself whose stencil stays inside it)
template <int S, bool HASW, bool CLIP, bool UNIT>
__device__ __forceinline__ void accumulate(double* acc, const float* r, int N, int ox, int oy, int oz, double inv,
                                           double mult) {
    double w0[S], w1[S], w2[S];
    int l0 = wrapN(Axis<S>::eval((double)r[0] * inv, w0), N) - ox;
    int l1 = wrapN(Axis<S>::eval((double)r[1] * inv, w1), N) - oy;
    int l2 = wrapN(Axis<S>::eval((double)r[2] * inv, w2), N) - oz;
    // a particle of the lower neighbour across the periodic boundary
    if (l0 >= TILE) l0 -= N;
    if (l1 >= TILE) l1 -= N;
    if (l2 >= TILE) l2 -= N;
    const double ww = HASW ? (double)r[3] * mult : mult;
#pragma unroll
    for (int l = 0; l < S; ++l) {
        const int c0 = l0 + l;
        if (CLIP && (c0 < 0 || c0 >= TILE)) continue;
#pragma unroll
        for (int m = 0; m < S; ++m) {
            const int c1 = l1 + m;
            if (CLIP && (c1 < 0 || c1 >= TILE)) continue;
            const double wlm = UNIT ? w0[l] * w1[m] : w0[l] * w1[m] * ww;   // UNIT: no weights, mult == 1
            double* row = acc + (c0 * TILE + c1) * TROW;
#pragma unroll
            for (int q = 0; q < S; ++q) {
                const int c2 = l2 + q;
                if (!CLIP || (c2 >= 0 && c2 < TILE)) atomicAdd(row + c2, wlm * w2[q]);
            }
        }
    }
}

// accumulate() with the 1-D weights and their products in float32 (Axis32); each contribution is
// added to the float64 LDS cell
template <int S, bool HASW, bool CLIP, bool UNIT>
__device__ __forceinline__ void accumulate32(double* acc, const float* r, int N, int ox, int oy, int oz, double inv,
                                             float mult) {
    float w0[S], w1[S], w2[S];
    int l0 = wrapN(Axis32<S>::eval((double)r[0] * inv, w0), N) - ox;
    int l1 = wrapN(Axis32<S>::eval((double)r[1] * inv, w1), N) - oy;
    int l2 = wrapN(Axis32<S>::eval((double)r[2] * inv, w2), N) - oz;
    if (l0 >= TILE) l0 -= N;
    if (l1 >= TILE) l1 -= N;
    if (l2 >= TILE) l2 -= N;
    const float ww = HASW ? r[3] * mult : mult;
#pragma unroll
    for (int l = 0; l < S; ++l) {
        const int c0 = l0 + l;
        if (CLIP && (c0 < 0 || c0 >= TILE)) continue;
#pragma unroll
        for (int m = 0; m < S; ++m) {
            const int c1 = l1 + m;
            if (CLIP && (c1 < 0 || c1 >= TILE)) continue;
            const float wlm = UNIT ? w0[l] * w1[m] : w0[l] * w1[m] * ww;
            double* row = acc + (c0 * TILE + c1) * TROW;
#pragma unroll
            for (int q = 0; q < S; ++q) {
                const int c2 = l2 + q;
                if (!CLIP || (c2 >= 0 && c2 < TILE)) atomicAdd(row + c2, (double)(wlm * w2[q]));
            }
        }
    }
}

// ACC: ACC_F64 (float64 weights, float64 cell sums), ACC_RUN (the same, with register runs of the
// stencil block for dense CIC inputs) or ACC_F32W (float32 weights and products, float64
// cell sums).  Both run at the same speed on gfx950 (the kernel is bound by loads and LDS
// atomics, not by float64 arithmetic); a float32 LDS cell (ds_add_f32) was measured about six times
// slower than the float64 one (ds_add_f64) in this kernel, and a TwoSum-compensated float32 cell
// (returning atomics) twelve times slower, so the float64 cell is the accumulator of every mode.
enum { ACC_F64 = 0, ACC_F32W = 1, ACC_RUN = 2 };

int dense_acc(int acc, int scheme, int64_t n, int dims) {
    return acc == ACC_F64 && scheme == 2 && (double)n >= RUN_PPC * (double)dims * dims * dims ? ACC_RUN : acc;
}

// k_bin counts runs of equal keys once from AGG_PPC particles per cell up: spatially ordered inputs
// then give runs of tens of lanes per key, whose per-lane LDS atomics on one word serialise (512^3
// lattice particles onto 256^3: 1.68 -> 1.41 ms for the deposit); at one particle per cell 512^3 the
// shuffles cost more than they save (+2 %), and unordered inputs are unchanged either way
constexpr double AGG_PPC = 2.0;
int bin_agg(int64_t n, int dims) { return (double)n >= AGG_PPC * (double)dims * dims * dims ? 1 : 0; }

template <int S, typename G, bool HASW, int MODE, bool UNIT, int ACC>
__device__ __forceinline__ void tile_body(const TileArgs& a, Blk B) {
    constexpr int R = HASW ? 4 : 3;
    constexpr int NE = 72;   // 8 neighbours (d = 0 own) x (8 class segments + 1 overflow group)
    __shared__ double acc[TILE * TILE * TROW];
    __shared__ long long sbase[NE];
    __shared__ unsigned scnt[NE];
    __shared__ unsigned sstart[NE + 1];
    const int tid = threadIdx.x;
    // contiguous runs of tiles per XCD: the lower neighbours whose halo records a tile reads were
    // mostly processed just before on the same XCD
    const int nt = a.nt, ntl = nt * nt * nt, tper = (ntl + XCDS - 1) / XCDS;
    const int t = (int)(B.x % XCDS) * tper + (int)(B.x / XCDS);
    if (t >= ntl) return;
    const int tx = t / (nt * nt), ty = (t / nt) % nt, tz = t % nt;
    for (int i = tid; i < TILE * TILE * TROW; i += TT) acc[i] = 0.0;
    // entry e = d * 9 + c: class segment c (< 8) or the overflow group (c = 8) of the neighbour
    // t - d; a neighbour's particles reach this tile when their class contains d
    if (tid < NE) {
        const int d = tid / 9, c = tid % 9;
        const int dx = d & 1, dy = (d >> 1) & 1, dz = (d >> 2) & 1;
        const long long tn = ((long long)((tx - dx + nt) % nt) * nt + (ty - dy + nt) % nt) * nt + (tz - dz + nt) % nt;
        unsigned n = 0u;
        long long b = 0;
        if (c < 8) {
            if ((c & d) == d) {
                n = min(a.fill[tn * 8 + c], (unsigned)a.cap[c]);
                b = tn * a.CT + a.cofs[c];
            }
        } else {
            n = a.ocnt[tn];
            b = a.ooff[tn];
        }
        scnt[tid] = n;
        sbase[tid] = b;
    }
    __syncthreads();
    if (tid < 64) {
        // exclusive prefix of the 72 counts: lane l holds entries 2l and 2l+1
        const int e0 = 2 * tid, e1 = 2 * tid + 1;
        const unsigned c0 = e0 < NE ? scnt[e0] : 0u, c1 = e1 < NE ? scnt[e1] : 0u;
        unsigned v = c0 + c1;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned u = __shfl_up(v, o, 64);
            if (tid >= o) v += u;
        }
        const unsigned ex = v - c0 - c1;
        if (e0 < NE) sstart[e0] = ex;
        if (e1 < NE) sstart[e1] = ex + c0;
        if (e1 == NE - 1) sstart[NE] = v;
    }
    __syncthreads();
    const unsigned tot = sstart[NE];
    const int ox = tx * TILE, oy = ty * TILE, oz = tz * TILE;
    // RP records per thread per step; the records of the next step load while those of this step
    // accumulate.  The loads are unconditional (an exhausted slot reads record 0 and class word 0,
    // both always allocated) and their values are first used a step later, so only they stay in
    // flight across the accumulation.
    constexpr int RP = 2;
    int e = 0;
    float cr[RP][4], nr[RP][4];
    int ce[RP], ne[RP];
    unsigned cg[RP], ng[RP];
    auto fetch = [&](unsigned f) {
#pragma unroll
        for (int j = 0; j < RP; ++j) {
            const unsigned fj = f + j * TT;
            const bool valid = fj < tot;
            if (valid)
                while (fj >= sstart[e + 1]) ++e;
            const bool ovg = e % 9 == 8;
            const long long i = valid ? sbase[e] + (fj - sstart[e]) : 0;
            const float* q = (valid && ovg ? a.grouped : a.bucket) + i * R;
#pragma unroll
            for (int c = 0; c < R; ++c) nr[j][c] = q[c];
            ng[j] = a.gcls[valid && ovg ? i : 0];
            ne[j] = valid ? e : -1;
        }
    };
    auto advance = [&]() {
#pragma unroll
        for (int j = 0; j < RP; ++j) {
#pragma unroll
            for (int c = 0; c < 4; ++c) cr[j][c] = nr[j][c];
            ce[j] = ne[j];
            cg[j] = ng[j];
        }
    };
    // phase 1: the tile's own class-0 records (most of them): one contiguous segment whose stencils
    // stay inside the tile -- no entry search, no clipping, no overflow class word
    const unsigned n0 = scnt[0];
    const float* q0 = a.bucket + sbase[0] * R;
    if constexpr (S == 2 && ACC == ACC_RUN) {
        // Runs (dense inputs): a thread takes groups of RG consecutive records (consecutive particles
        // of the input, mostly neighbours along z) and keeps the float64 sums of its current S^3
        // stencil block in registers.  A record with the same lowest cell adds to them; one cell
        // further along z, the block's lower z face goes to LDS and the block moves up (CIC: 4 LDS
        // atomics instead of 8); any other record sends the whole block to LDS.  With several
        // particles per cell most records add in registers (512^3 particles onto 256^3: 0.95 ->
        // 0.79 ms); with about one per cell the strided group loads and the register blocks cost more
        // than the atomics they save, so the host picks this mode by particles per cell.
        constexpr int RG = 2;
        double blk[S][S][S];
        int b0 = -1, b1 = 0, b2 = 0;   // lowest cell of the block (b0 < 0: none yet)
        float gr[RG][4], gn[RG][4];
        auto fetchg = [&](unsigned g) {
#pragma unroll
            for (int k = 0; k < RG; ++k) {
                const unsigned fk = g * RG + k;
                const float* q = q0 + (size_t)(fk < n0 ? fk : 0u) * R;   // record 0 of the segment always exists
#pragma unroll
                for (int c = 0; c < R; ++c) gn[k][c] = q[c];
            }
        };
        auto all_out = [&]() {
#pragma unroll
            for (int l = 0; l < S; ++l)
#pragma unroll
                for (int m = 0; m < S; ++m) {
                    double* row = acc + ((b0 + l) * TILE + b1 + m) * TROW + b2;
#pragma unroll
                    for (int q = 0; q < S; ++q) atomicAdd(row + q, blk[l][m][q]);
                }
        };
        const unsigned ngr = (n0 + RG - 1) / RG;
        fetchg(tid);
        for (unsigned g = tid; g < ngr; g += TT) {
#pragma unroll
            for (int k = 0; k < RG; ++k)
#pragma unroll
                for (int c = 0; c < 4; ++c) gr[k][c] = gn[k][c];
            fetchg(g + TT);
#pragma unroll
            for (int k = 0; k < RG; ++k) {
                if (g * RG + k >= n0) break;
                double w0[S], w1[S], w2[S];
                int l0 = wrapN(Axis<S>::eval((double)gr[k][0] * a.inv, w0), a.N) - ox;
                int l1 = wrapN(Axis<S>::eval((double)gr[k][1] * a.inv, w1), a.N) - oy;
                int l2 = wrapN(Axis<S>::eval((double)gr[k][2] * a.inv, w2), a.N) - oz;
                if (l0 >= TILE) l0 -= a.N;
                if (l1 >= TILE) l1 -= a.N;
                if (l2 >= TILE) l2 -= a.N;
                if (l0 != b0 || l1 != b1 || l2 != b2) {
                    if (S == 2 && l0 == b0 && l1 == b1 && l2 == b2 + 1) {
#pragma unroll
                        for (int l = 0; l < S; ++l)
#pragma unroll
                            for (int m = 0; m < S; ++m) {
                                atomicAdd(acc + ((b0 + l) * TILE + b1 + m) * TROW + b2, blk[l][m][0]);
                                blk[l][m][0] = blk[l][m][S - 1];
                                blk[l][m][S - 1] = 0.0;
                            }
                    } else {
                        if (b0 >= 0) all_out();
#pragma unroll
                        for (int l = 0; l < S; ++l)
#pragma unroll
                            for (int m = 0; m < S; ++m)
#pragma unroll
                                for (int q = 0; q < S; ++q) blk[l][m][q] = 0.0;
                        b0 = l0;
                        b1 = l1;
                    }
                    b2 = l2;
                }
                const double ww = HASW ? (double)gr[k][3] * a.mult : a.mult;
#pragma unroll
                for (int l = 0; l < S; ++l)
#pragma unroll
                    for (int m = 0; m < S; ++m) {
                        const double wlm = UNIT ? w0[l] * w1[m] : w0[l] * w1[m] * ww;
#pragma unroll
                        for (int q = 0; q < S; ++q) blk[l][m][q] += wlm * w2[q];
                    }
            }
        }
        if (b0 >= 0) all_out();
    } else {
        auto fetch0 = [&](unsigned f) {
#pragma unroll
            for (int j = 0; j < RP; ++j) {
                const unsigned fj = f + j * TT;
                const float* q = q0 + (size_t)(fj < n0 ? fj : 0u) * R;   // record 0 of the segment always exists
#pragma unroll
                for (int c = 0; c < R; ++c) nr[j][c] = q[c];
            }
        };
        fetch0(tid);
        for (unsigned f = tid; f < n0; f += RP * TT) {
#pragma unroll
            for (int j = 0; j < RP; ++j)
#pragma unroll
                for (int c = 0; c < 4; ++c) cr[j][c] = nr[j][c];
            fetch0(f + RP * TT);
#pragma unroll
            for (int j = 0; j < RP; ++j) {
                if (f + j * TT >= n0) break;
                if (ACC == ACC_F32W)
                    accumulate32<S, HASW, false, UNIT>(acc, cr[j], a.N, ox, oy, oz, a.inv, (float)a.mult);
                else
                    accumulate<S, HASW, false, UNIT>(acc, cr[j], a.N, ox, oy, oz, a.inv, a.mult);
            }
        }
    }
    // phase 2: halo-class segments of this tile and its lower neighbours, overflow groups
    fetch(n0 + tid);
    advance();
    for (unsigned f = n0 + tid; f < tot; f += RP * TT) {
        fetch(f + RP * TT);
#pragma unroll
        for (int j = 0; j < RP; ++j) {
            const int ej = ce[j];
            // entry 0: the tile's own class-0 particles, whose stencils stay inside the tile;
            // an overflow group's record reaches this tile when its class contains the offset
            const unsigned d = (unsigned)(ej / 9);
            const bool ok = ej >= 0 && (ej % 9 < 8 || (cg[j] & d) == d);
            if (ACC == ACC_F32W) {
                if (ok) accumulate32<S, HASW, true, UNIT>(acc, cr[j], a.N, ox, oy, oz, a.inv, (float)a.mult);
            } else if (ok) {
                accumulate<S, HASW, true, UNIT>(acc, cr[j], a.N, ox, oy, oz, a.inv, a.mult);
            }
        }
        advance();
    }
    __syncthreads();
    G* grid = (G*)a.grid;
    const double am = (MODE == OUT_DELTA) ? (double)a.N * a.N * a.N / (a.dmass ? a.dmass[0] : a.mass) : 0.0;
    const long long N = a.N;
    for (int idx = tid; idx < TILE * TILE * TILE; idx += TT) {
        const int i = idx / (TILE * TILE), j = (idx / TILE) % TILE, k = idx % TILE;
        const long long g = (((long long)ox + i) * N + oy + j) * N + oz + k;
        put<G, MODE>(grid, g, acc[(i * TILE + j) * TROW + k], am);
    }
}

template <int S, typename G, bool HASW, int MODE, bool UNIT, int ACC>
__global__ __launch_bounds__(TT) void k_tile(TileArgs a) {
    tile_body<S, G, HASW, MODE, UNIT, ACC>(a, own_blk());
}
template <int S, typename G, bool HASW, int MODE, bool UNIT, int ACC>
__global__ __launch_bounds__(TT) void k_tile_b(const TileArgs* __restrict__ tab,
                                                  const unsigned* __restrict__ start, int nj) {
    PYL_BATCH_BLK(start, nj);
    tile_body<S, G, HASW, MODE, UNIT, ACC>(tab[bj_], blk_);
}

// ------------------------------------------------------------------ presort for unordered inputs
// Particles in no spatial order give k_bin one key per particle: one global atomic each and 12-byte
// scattered records.  The presort first groups them by supertile (tps consecutive tiles, at most
// SUPMAX groups) with an exact count, a scan and a scatter whose per-step chunks are contiguous; k_bin
// then reads the grouped copy, where neighbouring particles share keys as in an ordered input.
constexpr int PT = 512;   // presort kernels: threads per workgroup
constexpr int PKP = 8;    // k_sup_scatter: particles per thread per step


__device__ __forceinline__ void sup_load(const SupArgs& a, long long p, float& x, float& y, float& z) {
    float c[3] = {a.pos[3 * p], a.pos[3 * p + 1], a.pos[3 * p + 2]};
    if (a.rsd.vel) c[a.rsd.axis] = rsd_shift(c[a.rsd.axis], a.rsd.vel[3 * p + a.rsd.axis], a.rsd.fac, a.rsd.box);
    x = c[0];
    y = c[1];
    z = c[2];
}

// supertile of a particle.  The grouping only has to be deterministic (the count and the scatter
// must agree) and spatially coherent, not equal to k_bin's tiles, so float32 cells clamped into
// the grid do (NaN goes to cell 0).
__device__ __forceinline__ int sup_cell(float x, float inv, int N) {
    return (int)fminf(fmaxf(x * inv, 0.0f), (float)(N - 1));
}
template <int S>
__device__ __forceinline__ int sup_of(const SupArgs& a, float x, float y, float z) {
    const float inv = (float)a.inv;
    const int bx = sup_cell(x, inv, a.N), by = sup_cell(y, inv, a.N), bz = sup_cell(z, inv, a.N);
    return (((bx / TILE) * a.nt + by / TILE) * a.nt + bz / TILE) / a.tps;
}

template <int S>
__device__ __forceinline__ void sup_count_body(const SupArgs& a, Blk B) {
    __shared__ unsigned h[SUPMAX];
    for (int i = threadIdx.x; i < a.nsup; i += PT) h[i] = 0u;
    __syncthreads();
    for (long long p = (long long)B.x * PT + threadIdx.x; p < a.n; p += (long long)B.gx * PT) {
        float x, y, z;
        sup_load(a, p, x, y, z);
        atomicAdd(&h[sup_of<S>(a, x, y, z)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < a.nsup; i += PT)
        if (h[i]) atomicAdd(&a.cnt[i], h[i]);
}

template <int S>
__global__ __launch_bounds__(PT) void k_sup_count(SupArgs a) { sup_count_body<S>(a, own_blk()); }
template <int S>
__global__ __launch_bounds__(PT) void k_sup_count_b(const SupArgs* __restrict__ tab,
                                                       const unsigned* __restrict__ start, int nj) {
    PYL_BATCH_BLK(start, nj);
    sup_count_body<S>(tab[bj_], blk_);
}

// exclusive scan of the nsup <= SUPMAX counts (one workgroup of SUPMAX threads)
__device__ __forceinline__ void sup_scan_body(const unsigned* __restrict__ cnt, unsigned* __restrict__ off, int nsup) {
    __shared__ unsigned wsum[SUPMAX / 64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const unsigned c = tid < nsup ? cnt[tid] : 0u;
    unsigned v = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) wsum[wv] = v;
    __syncthreads();
    unsigned before = 0u;
    for (int j = 0; j < wv; ++j) before += wsum[j];
    if (tid < nsup) off[tid] = before + v - c;
}
__global__ __launch_bounds__(SUPMAX) void k_sup_scan(const unsigned* __restrict__ cnt, unsigned* __restrict__ off,
                                                     int nsup) {
    sup_scan_body(cnt, off, nsup);
}
// batched: one workgroup per job
__global__ __launch_bounds__(SUPMAX) void k_sup_scan_b(const SupArgs* __restrict__ tab) {
    const SupArgs& a = tab[blockIdx.x];
    sup_scan_body(a.cnt, a.off, a.nsup);
}

template <int S, bool HASW>
__device__ __forceinline__ void sup_scatter_body(const SupArgs& a, Blk B) {
    constexpr long long PS = (long long)PT * PKP;
    __shared__ unsigned lc[SUPMAX], lb[SUPMAX];
    const int tid = threadIdx.x;
    for (int i = tid; i < a.nsup; i += PT) lc[i] = 0u;
    __syncthreads();
    const long long nstep = (a.n + PS - 1) / PS;
    for (long long step = B.x; step < nstep; step += B.gx) {
        float x[PKP], y[PKP], z[PKP], w[PKP];
        int sp[PKP];
        unsigned rank[PKP];
#pragma unroll
        for (int k = 0; k < PKP; ++k) {
            const long long p = step * PS + k * PT + tid;
            sp[k] = -1;
            rank[k] = 0u;
            if (p < a.n) {
                sup_load(a, p, x[k], y[k], z[k]);
                w[k] = HASW ? a.W[p] : 0.f;
                sp[k] = sup_of<S>(a, x[k], y[k], z[k]);
                rank[k] = atomicAdd(&lc[sp[k]], 1u);
            }
        }
        __syncthreads();
        // one reservation per supertile present in the step; the counters are cleared for the next
        // step, whose insertions every thread makes only after its stores below
        for (int i = tid; i < a.nsup; i += PT) {
            const unsigned c = lc[i];
            if (c) {
                lb[i] = a.off[i] + atomicAdd(&a.cur[i], c);
                lc[i] = 0u;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PKP; ++k) {
            if (sp[k] < 0) continue;
            const long long d = (long long)lb[sp[k]] + rank[k];
            if (d >= a.n) continue;   // unreachable: the count pass gave every particle its slot
            a.spos[3 * d] = x[k];
            a.spos[3 * d + 1] = y[k];
            a.spos[3 * d + 2] = z[k];
            if (HASW) a.sw[d] = w[k];
        }
    }
}

template <int S, bool HASW>
__global__ __launch_bounds__(PT) void k_sup_scatter(SupArgs a) { sup_scatter_body<S, HASW>(a, own_blk()); }
template <int S, bool HASW>
__global__ __launch_bounds__(PT) void k_sup_scatter_b(const SupArgs* __restrict__ tab,
                                                         const unsigned* __restrict__ start, int nj) {
    PYL_BATCH_BLK(start, nj);
    sup_scatter_body<S, HASW>(tab[bj_], blk_);
}

// ------------------------------------------------------------------ host side
bool binned_supported(int axes, int dims) { return axes == 3 && dims % TILE == 0 && dims >= 2 * TILE; }

// Bucket layout of one deposit: class segment capacities of twice the expected count of a uniform
// distribution plus slack, and the bytes of the counters, buckets and overflow area.
struct Layout {
    int nt, cap[8], cofs[8];
    long long ntiles, CT;
    size_t ncnt;   // counters: fill (8 ntiles) | ocnt | ocur | ooff (ntiles + 1) | ov_count
    size_t cnt_bytes, bucket_bytes, ov_bytes;
};

static int layout_of(int64_t n, int dims, int S, bool hasw, Layout* L) {
    const int R = hasw ? 4 : 3;
    L->nt = dims / TILE;
    L->ntiles = (long long)L->nt * L->nt * L->nt;
    const double avg = (double)n / (double)L->ntiles;
    const double ph = (double)(S - 1) / TILE, p0 = 1.0 - ph;
    L->CT = 0;
    for (int c = 0; c < 8; ++c) {
        const double f = ((c & 1) ? ph : p0) * ((c & 2) ? ph : p0) * ((c & 4) ? ph : p0);
        const long long v = f > 0.0 ? ((long long)(2.0 * avg * f) + 32 + 15) / 16 * 16 : 0;
        PYL_CHECK(v < (1LL << 30), E_UNSUPPORTED, "binned deposit: too many particles per tile");
        L->cap[c] = (int)v;
        L->cofs[c] = (int)L->CT;
        L->CT += v;
    }
    L->ncnt = 11 * (size_t)L->ntiles + 2;
    L->cnt_bytes = sizeof(unsigned) * L->ncnt + 64;
    L->bucket_bytes = sizeof(float) * (size_t)L->ntiles * L->CT * R;
    // overflow: ov_idx (8n) | ov_key (4n) | gcls (4n) | grouped (4Rn)
    L->ov_bytes = (size_t)n * (16 + 4 * R) + 256;
    return OK;
}

template <int S, typename G, bool HASW, int MODE>
static int run_binned(const float* pos, const float* W, int64_t n, int dims, double BoxSize, double mult, void* grid,
                      const double* dmass, double mass, hipStream_t st, const Rsd& rsd, int acc) {
    int rc = OK;
    Layout L;
    rc = layout_of(n, dims, S, HASW, &L);
    if (rc) return rc;
    const int nt = L.nt;
    const long long ntiles = L.ntiles, CT = L.CT;
    const int *cap = L.cap, *cofs = L.cofs;
    const size_t ncnt = L.ncnt;
    unsigned* cnt = (unsigned*)workspace(WS_BINS, L.cnt_bytes, &rc);
    if (rc) return rc;
    unsigned *fill = cnt, *ocnt = cnt + 8 * ntiles, *ocur = cnt + 9 * ntiles, *ooff = cnt + 10 * ntiles,
             *ov_count = cnt + 11 * ntiles + 1;
    float* bucket = (float*)workspace(WS_AUX, L.bucket_bytes, &rc);
    if (rc) return rc;
    char* ov = (char*)workspace(WS_AUX2, L.ov_bytes, &rc);
    if (rc) return rc;
    long long* ov_idx = (long long*)ov;
    unsigned* ov_key = (unsigned*)(ov + 8 * (size_t)n);
    unsigned* gcls = (unsigned*)(ov + 12 * (size_t)n);
    float* grouped = (float*)(ov + 16 * (size_t)n + 64);
    rc = zero_fill(cnt, sizeof(unsigned) * ncnt, st);
    if (rc) return rc;
    {
        StageTimer tm("deposit_bin", st);
        BinArgs ba{pos, W, (long long)n, dims, nt, dims / BoxSize, fill, bucket, CT, {}, {}, ov_count, ov_idx, ov_key, rsd};
        ba.agg = bin_agg(n, dims);
        for (int c = 0; c < 8; ++c) {
            ba.cofs[c] = cofs[c];
            ba.cap[c] = cap[c];
        }
        long long blocks = (n + (long long)BT * KP - 1) / ((long long)BT * KP);
        if (blocks > 2048) blocks = 2048;
        blocks = (blocks + XCDS - 1) / XCDS * XCDS;
        if (ba.agg)
            hipLaunchKernelGGL((k_bin<S, HASW, 1>), dim3((unsigned)blocks), dim3(BT), 0, st, ba);
        else
            hipLaunchKernelGGL((k_bin<S, HASW, 0>), dim3((unsigned)blocks), dim3(BT), 0, st, ba);
        PYL_HIP(hipGetLastError());
    }
    {
        StageTimer tm("deposit_overflow", st);
        const OvArgs oa{ov_count, ov_idx, ov_key, ocnt, ooff, ocur, pos, W, grouped, gcls, rsd, (int)ntiles};
        hipLaunchKernelGGL(k_ov_count, dim3(256), dim3(256), 0, st, oa);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, oa);
        hipLaunchKernelGGL((k_ov_scatter<HASW>), dim3(256), dim3(256), 0, st, oa);
        PYL_HIP(hipGetLastError());
    }
    {
        StageTimer tm("deposit_tiles", st);
        TileArgs ta{bucket, fill, CT, {}, {}, grouped, gcls, ooff, ocnt, dims, nt, dims / BoxSize, mult, grid, dmass,
                    mass};
        for (int c = 0; c < 8; ++c) {
            ta.cofs[c] = cofs[c];
            ta.cap[c] = cap[c];
        }
        const dim3 tg((unsigned)((ntiles + XCDS - 1) / XCDS * XCDS));
        const bool unit = !HASW && mult == 1.0;
#define PYL_TILE_LAUNCH(U, A) hipLaunchKernelGGL((k_tile<S, G, HASW, MODE, U, A>), tg, dim3(TT), 0, st, ta)
        if constexpr (S == 2) {
            if (acc == ACC_RUN) {
                if (unit) PYL_TILE_LAUNCH(true, ACC_RUN); else PYL_TILE_LAUNCH(false, ACC_RUN);
            }
        }
        if (acc == ACC_RUN) {
        } else if (acc == ACC_F32W) {
            if (unit) PYL_TILE_LAUNCH(true, ACC_F32W); else PYL_TILE_LAUNCH(false, ACC_F32W);
        } else {
            if (unit) PYL_TILE_LAUNCH(true, ACC_F64); else PYL_TILE_LAUNCH(false, ACC_F64);
        }
#undef PYL_TILE_LAUNCH
        PYL_HIP(hipGetLastError());
    }
    return OK;
}

template <int S, typename G, int MODE>
static int dispatch_w(const float* pos, const float* W, int64_t n, int dims, double BoxSize, double mult, void* grid,
                      const double* dmass, double mass, hipStream_t st, const Rsd& rsd, int acc) {
    return W ? run_binned<S, G, true, MODE>(pos, W, n, dims, BoxSize, mult, grid, dmass, mass, st, rsd, acc)
             : run_binned<S, G, false, MODE>(pos, W, n, dims, BoxSize, mult, grid, dmass, mass, st, rsd, acc);
}

template <typename G, int MODE>
static int dispatch_s(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, double mult,
                      void* grid, const double* dmass, double mass, hipStream_t st, const Rsd& rsd, int acc) {
    switch (scheme) {
        case 1: return dispatch_w<1, G, MODE>(pos, W, n, dims, BoxSize, mult, grid, dmass, mass, st, rsd, acc);
        case 2: return dispatch_w<2, G, MODE>(pos, W, n, dims, BoxSize, mult, grid, dmass, mass, st, rsd, acc);
        case 3: return dispatch_w<3, G, MODE>(pos, W, n, dims, BoxSize, mult, grid, dmass, mass, st, rsd, acc);
        case 4: return dispatch_w<4, G, MODE>(pos, W, n, dims, BoxSize, mult, grid, dmass, mass, st, rsd, acc);
        default: set_error("unknown scheme %d", scheme); return E_ARG;
    }
}

template <int S>
static void launch_presort(const SupArgs& a, bool hasw, unsigned blocks, unsigned sblocks, hipStream_t st) {
    hipLaunchKernelGGL(k_sup_count<S>, dim3(blocks), dim3(PT), 0, st, a);
    hipLaunchKernelGGL(k_sup_scan, dim3(1), dim3(SUPMAX), 0, st, a.cnt, a.off, a.nsup);
    if (hasw)
        hipLaunchKernelGGL((k_sup_scatter<S, true>), dim3(sblocks), dim3(PT), 0, st, a);
    else
        hipLaunchKernelGGL((k_sup_scatter<S, false>), dim3(sblocks), dim3(PT), 0, st, a);
}

// groups the particles by supertile into library workspaces (*spos, *sw); the redshift-space shift
// is applied on the way, so the grouped copy is deposited without it
static int presort(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, const Rsd& rsd,
                   hipStream_t st, const float** spos, const float** sw) {
    int rc = OK;
    const int nt = dims / TILE;
    const long long ntiles = (long long)nt * nt * nt;
    // supertiles of whole tile columns (nt tiles along z), about 128 of them and never more than SUPMAX
    long long tps = nt * std::max(1LL, (ntiles + 128LL * nt - 1) / (128LL * nt));
    while ((ntiles + tps - 1) / tps > SUPMAX) tps += nt;
    const int nsup = (int)((ntiles + tps - 1) / tps);
    unsigned* cnt = (unsigned*)workspace(WS_SUP, sizeof(unsigned) * 3 * SUPMAX, &rc);
    if (rc) return rc;
    float* sp = (float*)workspace(WS_STAGE, sizeof(float) * (size_t)n * (W ? 4 : 3) + 64, &rc);
    if (rc) return rc;
    rc = zero_fill(cnt, sizeof(unsigned) * 3 * SUPMAX, st);
    if (rc) return rc;
    SupArgs a{pos, W, (long long)n, dims, nt, (int)tps, nsup, dims / BoxSize, rsd, cnt, cnt + SUPMAX,
              cnt + 2 * SUPMAX, sp, W ? sp + 3 * n : nullptr};
    long long blocks = (n + PT - 1) / PT;
    if (blocks > 2048) blocks = 2048;
    long long sblocks = (n + (long long)PT * PKP - 1) / ((long long)PT * PKP);
    if (sblocks > 2048) sblocks = 2048;
    StageTimer tm("deposit_presort", st);
    switch (scheme) {
        case 1: launch_presort<1>(a, W != nullptr, (unsigned)blocks, (unsigned)sblocks, st); break;
        case 2: launch_presort<2>(a, W != nullptr, (unsigned)blocks, (unsigned)sblocks, st); break;
        case 3: launch_presort<3>(a, W != nullptr, (unsigned)blocks, (unsigned)sblocks, st); break;
        case 4: launch_presort<4>(a, W != nullptr, (unsigned)blocks, (unsigned)sblocks, st); break;
        default: set_error("unknown scheme %d", scheme); return E_ARG;
    }
    PYL_HIP(hipGetLastError());
    *spos = sp;
    *sw = a.sw;
    return OK;
}

// mode: 0 accumulate, 1 overwrite, 2 overdensity (mass = dmass[0] if dmass else mass); acc: cell
// accumulation of a float32 grid (ACC_F64 / ACC_F32W; float64 grids always use ACC_F64)
int deposit_binned(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, bool f64,
                   int mode, double mult, void* grid, const double* dmass, double mass, hipStream_t st,
                   const Rsd& rsd_in, bool sorted, int acc) {
    PYL_CHECK(!rsd_in.vel || (rsd_in.axis >= 0 && rsd_in.axis <= 2), E_ARG, "redshift-space axis must be 0, 1 or 2");
    PYL_CHECK(binned_supported(3, dims), E_UNSUPPORTED, "binned deposit needs 3-D grids with dims %% %d == 0", TILE);
    PYL_CHECK(n > 0, E_ARG, "binned deposit needs particles");
    PYL_CHECK(n < (1LL << 31), E_UNSUPPORTED, "binned deposit: at most 2^31 particles per call");
    Rsd rsd = rsd_in;
    if (sorted) {
        const float *sp = nullptr, *sw = nullptr;
        int rc = presort(pos, W, n, dims, BoxSize, scheme, rsd_in, st, &sp, &sw);
        if (rc) return rc;
        pos = sp;
        W = sw;
        rsd = Rsd();
    }
    if (f64) {
        const int a64 = dense_acc(ACC_F64, scheme, n, dims);
        if (mode == OUT_ACC) return dispatch_s<double, OUT_ACC>(pos, W, n, dims, BoxSize, scheme, mult, grid, dmass, mass, st, rsd, a64);
        if (mode == OUT_SET) return dispatch_s<double, OUT_SET>(pos, W, n, dims, BoxSize, scheme, mult, grid, dmass, mass, st, rsd, a64);
        return dispatch_s<double, OUT_DELTA>(pos, W, n, dims, BoxSize, scheme, mult, grid, dmass, mass, st, rsd, a64);
    }
    acc = dense_acc(acc, scheme, n, dims);
    if (mode == OUT_ACC) return dispatch_s<float, OUT_ACC>(pos, W, n, dims, BoxSize, scheme, mult, grid, dmass, mass, st, rsd, acc);
    if (mode == OUT_SET) return dispatch_s<float, OUT_SET>(pos, W, n, dims, BoxSize, scheme, mult, grid, dmass, mass, st, rsd, acc);
    return dispatch_s<float, OUT_DELTA>(pos, W, n, dims, BoxSize, scheme, mult, grid, dmass, mass, st, rsd, acc);
}

// ------------------------------------------------------------------ batched deposit
// Jobs of a batch (pyl_graph_create / pyl_pk_batch) run each stage of the deposit as ONE launch for
// all jobs of a class (same scheme, weights, grid type, accumulation and presort): the arguments of
// job j sit in device tables and its blocks are the range [start[j], start[j+1]) of the launch.
static size_t al256(size_t b) { return (b + 255) / 256 * 256; }

static void sup_geometry(int dims, int* tps, int* nsup) {
    const int nt = dims / TILE;
    const long long ntiles = (long long)nt * nt * nt;
    long long t = nt * std::max(1LL, (ntiles + 128LL * nt - 1) / (128LL * nt));
    while ((ntiles + t - 1) / t > SUPMAX) t += nt;
    *tps = (int)t;
    *nsup = (int)((ntiles + t - 1) / t);
}

size_t binned_job_bytes(int64_t n, int dims, int scheme, bool hasw, bool sorted) {
    Layout L;
    if (!binned_supported(3, dims) || layout_of(n, dims, scheme, hasw, &L) != OK) return 0;
    size_t b = al256(L.cnt_bytes) + al256(L.bucket_bytes) + al256(L.ov_bytes);
    if (sorted) b += al256(sizeof(unsigned) * 3 * SUPMAX) + al256(sizeof(float) * (size_t)n * (hasw ? 4 : 3) + 64);
    return b;
}

int binned_job_plan(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, void* grid,
                    const double* dmass, double mass, bool sorted, char* mem, BinnedJob* j) {
    PYL_CHECK(binned_supported(3, dims), E_UNSUPPORTED, "binned deposit needs 3-D grids with dims %% %d == 0", TILE);
    PYL_CHECK(n > 0 && n < (1LL << 31), E_UNSUPPORTED, "binned deposit: 1 to 2^31 - 1 particles per job");
    const bool hasw = W != nullptr;
    const int R = hasw ? 4 : 3;
    Layout L;
    int rc = layout_of(n, dims, scheme, hasw, &L);
    if (rc) return rc;
    char* p = mem;
    unsigned* cnt = (unsigned*)p;
    p += al256(L.cnt_bytes);
    float* bucket = (float*)p;
    p += al256(L.bucket_bytes);
    char* ov = p;
    p += al256(L.ov_bytes);
    const long long ntiles = L.ntiles;
    unsigned *fill = cnt, *ocnt = cnt + 8 * ntiles, *ocur = cnt + 9 * ntiles, *ooff = cnt + 10 * ntiles,
             *ov_count = cnt + 11 * ntiles + 1;
    long long* ov_idx = (long long*)ov;
    unsigned* ov_key = (unsigned*)(ov + 8 * (size_t)n);
    unsigned* gcls = (unsigned*)(ov + 12 * (size_t)n);
    float* grouped = (float*)(ov + 16 * (size_t)n + 64);
    const double inv = dims / BoxSize;
    const float* bpos = pos;
    const float* bW = W;
    j->sup = SupArgs{};
    j->sup_blocks = j->sup_sblocks = 0;
    j->scnt = nullptr;
    j->scnt_bytes = 0;
    if (sorted) {
        unsigned* sc = (unsigned*)p;
        p += al256(sizeof(unsigned) * 3 * SUPMAX);
        float* sp = (float*)p;
        int tps = 0, nsup = 0;
        sup_geometry(dims, &tps, &nsup);
        j->sup = SupArgs{pos, W, (long long)n, dims, L.nt, tps, nsup, inv, Rsd(), sc, sc + SUPMAX, sc + 2 * SUPMAX,
                         sp, hasw ? sp + 3 * n : nullptr};
        j->sup_blocks = (unsigned)std::min<long long>(2048, (n + PT - 1) / PT);
        j->sup_sblocks = (unsigned)std::min<long long>(2048, (n + (long long)PT * PKP - 1) / ((long long)PT * PKP));
        j->scnt = sc;
        j->scnt_bytes = sizeof(unsigned) * 3 * SUPMAX;
        bpos = sp;
        bW = j->sup.sw;
    }
    (void)R;
    j->bin = BinArgs{bpos, bW, (long long)n, dims, L.nt, inv, fill, bucket, L.CT, {}, {}, ov_count, ov_idx, ov_key,
                     Rsd()};
    j->bin.agg = bin_agg(n, dims);
    j->tile = TileArgs{bucket, fill, L.CT, {}, {}, grouped, gcls, ooff, ocnt, dims, L.nt, inv, 1.0, grid, dmass, mass};
    for (int c = 0; c < 8; ++c) {
        j->bin.cofs[c] = j->tile.cofs[c] = L.cofs[c];
        j->bin.cap[c] = j->tile.cap[c] = L.cap[c];
    }
    long long bb = (n + (long long)BT * KP - 1) / ((long long)BT * KP);
    if (bb > 2048) bb = 2048;
    j->bin_blocks = (unsigned)((bb + XCDS - 1) / XCDS * XCDS);
    j->ov = OvArgs{ov_count, ov_idx, ov_key, ocnt, ooff, ocur, bpos, bW, grouped, gcls, Rsd(), (int)ntiles};
    j->tile_blocks = (unsigned)((ntiles + XCDS - 1) / XCDS * XCDS);
    j->cnt = cnt;
    j->cnt_bytes = sizeof(unsigned) * L.ncnt;
    return OK;
}

template <int S, bool HASW, typename G, int ACC>
static int batch_launch_s(bool sorted, const BinnedTables& t, hipStream_t st) {
    const unsigned nj = (unsigned)t.nj;
    if (sorted) {
        StageTimer tm("deposit_presort", st);
        hipLaunchKernelGGL((k_sup_count_b<S>), dim3(t.sup_total), dim3(PT), 0, st, t.sup, t.sup_start, t.nj);
        hipLaunchKernelGGL(k_sup_scan_b, dim3(nj), dim3(SUPMAX), 0, st, t.sup);
        hipLaunchKernelGGL((k_sup_scatter_b<S, HASW>), dim3(t.ssc_total), dim3(PT), 0, st, t.sup, t.ssc_start, t.nj);
        PYL_HIP(hipGetLastError());
    }
    {
        StageTimer tm("deposit_bin", st);
        hipLaunchKernelGGL((k_bin_b<S, HASW>), dim3(t.bin_total), dim3(BT), 0, st, t.bin, t.bin_start, t.nj);
        PYL_HIP(hipGetLastError());
    }
    {
        StageTimer tm("deposit_overflow", st);
        hipLaunchKernelGGL(k_ov_count_b, dim3(256, nj), dim3(256), 0, st, t.ov);
        hipLaunchKernelGGL(k_scan_b, dim3(nj), dim3(1024), 0, st, t.ov);
        hipLaunchKernelGGL((k_ov_scatter_b<HASW>), dim3(256, nj), dim3(256), 0, st, t.ov);
        PYL_HIP(hipGetLastError());
    }
    {
        StageTimer tm("deposit_tiles", st);
        hipLaunchKernelGGL((k_tile_b<S, G, HASW, OUT_DELTA, !HASW, ACC>), dim3(t.tile_total), dim3(TT), 0, st, t.tile,
                           t.tile_start, t.nj);
        PYL_HIP(hipGetLastError());
    }
    return OK;
}

template <int S>
static int batch_launch_w(bool hasw, bool f64, int acc, bool sorted, const BinnedTables& t, hipStream_t st) {
    if constexpr (S == 2) {
        if (acc == ACC_RUN)
            return f64 ? (hasw ? batch_launch_s<S, true, double, ACC_RUN>(sorted, t, st)
                               : batch_launch_s<S, false, double, ACC_RUN>(sorted, t, st))
                       : (hasw ? batch_launch_s<S, true, float, ACC_RUN>(sorted, t, st)
                               : batch_launch_s<S, false, float, ACC_RUN>(sorted, t, st));
    }
    if (f64) return hasw ? batch_launch_s<S, true, double, ACC_F64>(sorted, t, st)
                         : batch_launch_s<S, false, double, ACC_F64>(sorted, t, st);
    if (acc == ACC_F32W) return hasw ? batch_launch_s<S, true, float, ACC_F32W>(sorted, t, st)
                                     : batch_launch_s<S, false, float, ACC_F32W>(sorted, t, st);
    return hasw ? batch_launch_s<S, true, float, ACC_F64>(sorted, t, st)
                : batch_launch_s<S, false, float, ACC_F64>(sorted, t, st);
}

int binned_batch_launch(int scheme, bool hasw, bool f64, int acc, bool sorted, const BinnedTables& t, hipStream_t st) {
    if (t.nj <= 0) return OK;
    switch (scheme) {
        case 1: return batch_launch_w<1>(hasw, f64, acc, sorted, t, st);
        case 2: return batch_launch_w<2>(hasw, f64, acc, sorted, t, st);
        case 3: return batch_launch_w<3>(hasw, f64, acc, sorted, t, st);
        case 4: return batch_launch_w<4>(hasw, f64, acc, sorted, t, st);
        default: set_error("unknown scheme %d", scheme); return E_ARG;
    }
}

}  // namespace pyl
