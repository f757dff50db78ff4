// LDS-staged Stockham FFT engine for gfx950 (64-wide wavefronts).
//
// A workgroup holds NL independent lines of length N in LDS; element (line l, index j)
// lives at s[j*ES + l] (ES >= NL), so consecutive lanes touch consecutive addresses.
// Every radix-RAD stage is one read phase (each thread pulls the RAD inputs of its
// butterflies into VGPRs), one barrier, and one write phase into the autosorted
// Stockham positions, so a single LDS buffer suffices.  The factorisation of N into
// radices 8/4/2/3/5 is compile-time, so every index is a constant-divisor expression.
//
// Forward transform: X[k] = sum_j x[j] exp(-2 pi i jk/N).  The inverse is obtained by
// conjugating on load and on store.  Twiddles come from a table tw[t] = exp(-2 pi i t/NT)
// with NT = N * TWSTEP, staged in LDS by the caller.
#pragma once
#include "common.hpp"

namespace pyl {

struct Radices {
    int n;
    int r[16];
};

constexpr Radices make_radices(int N) {
    Radices p{0, {0}};
    int m = N;
    // prefer radix 8; if N has 2^(3a+1) use one radix-2, 2^(3a+2) one radix-4
    int twos = 0;
    while (m % 2 == 0) { m /= 2; ++twos; }
    while (twos >= 3) { p.r[p.n++] = 8; twos -= 3; }
    if (twos == 2) p.r[p.n++] = 4;
    if (twos == 1) p.r[p.n++] = 2;
    while (m % 3 == 0) { p.r[p.n++] = 3; m /= 3; }
    while (m % 5 == 0) { p.r[p.n++] = 5; m /= 5; }
    if (m != 1) p.n = -1;  // unsupported length
    return p;
}

constexpr bool fft_len_ok(int N) { return N >= 2 && make_radices(N).n > 0; }

// ------------------------------------------------------------------ butterflies
template <typename R, int RAD> struct Dft;

template <typename R> struct Dft<R, 2> {
    static __device__ __forceinline__ void run(cplx<R>* a) {
        cplx<R> t = a[0] - a[1];
        a[0] = a[0] + a[1];
        a[1] = t;
    }
};

template <typename R> struct Dft<R, 4> {
    static __device__ __forceinline__ void run(cplx<R>* a) {
        cplx<R> t0 = a[0] + a[2], t1 = a[0] - a[2];
        cplx<R> t2 = a[1] + a[3], t3 = mul_mi(a[1] - a[3]);
        a[0] = t0 + t2;
        a[2] = t0 - t2;
        a[1] = t1 + t3;
        a[3] = t1 - t3;
    }
};

template <typename R> struct Dft<R, 8> {
    static __device__ __forceinline__ void run(cplx<R>* a) {
        const R h = (R)0.70710678118654752440084436210485;
        cplx<R> e[4] = {a[0], a[2], a[4], a[6]};
        cplx<R> o[4] = {a[1], a[3], a[5], a[7]};
        Dft<R, 4>::run(e);
        Dft<R, 4>::run(o);
        // W8^1 = (1 - i)/sqrt2, W8^2 = -i, W8^3 = (-1 - i)/sqrt2
        cplx<R> o1 = mk((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
        cplx<R> o2 = mul_mi(o[2]);
        cplx<R> o3 = mk((o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h);
        a[0] = e[0] + o[0];
        a[4] = e[0] - o[0];
        a[1] = e[1] + o1;
        a[5] = e[1] - o1;
        a[2] = e[2] + o2;
        a[6] = e[2] - o2;
        a[3] = e[3] + o3;
        a[7] = e[3] - o3;
    }
};

template <typename R> struct Dft<R, 3> {
    static __device__ __forceinline__ void run(cplx<R>* a) {
        const R s3 = (R)0.86602540378443864676372317075294;  // sin(2pi/3)
        cplx<R> s = a[1] + a[2], d = a[1] - a[2];
        cplx<R> m = mk(a[0].x - (R)0.5 * s.x, a[0].y - (R)0.5 * s.y);
        cplx<R> r = mul_mi(scale(d, s3));
        a[0] = a[0] + s;
        a[1] = m + r;
        a[2] = m - r;
    }
};

template <typename R> struct Dft<R, 5> {
    static __device__ __forceinline__ void run(cplx<R>* a) {
        const R c1 = (R)0.30901699437494742410229341718282;   // cos(2pi/5)
        const R c2 = (R)-0.80901699437494742410229341718282;  // cos(4pi/5)
        const R s1 = (R)0.95105651629515357211643933337938;   // sin(2pi/5)
        const R s2 = (R)0.58778525229247312916870595463907;   // sin(4pi/5)
        cplx<R> p1 = a[1] + a[4], d1 = a[1] - a[4];
        cplx<R> p2 = a[2] + a[3], d2 = a[2] - a[3];
        cplx<R> m1 = mk(a[0].x + c1 * p1.x + c2 * p2.x, a[0].y + c1 * p1.y + c2 * p2.y);
        cplx<R> m2 = mk(a[0].x + c2 * p1.x + c1 * p2.x, a[0].y + c2 * p1.y + c1 * p2.y);
        cplx<R> r1 = mul_mi(mk(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
        cplx<R> r2 = mul_mi(mk(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
        a[0] = a[0] + p1 + p2;
        a[1] = m1 + r1;
        a[4] = m1 - r1;
        a[2] = m2 + r2;
        a[3] = m2 - r2;
    }
};

// ------------------------------------------------------------------ one stage
template <typename R, int N, int NL, int ES, int T, int RAD, int NS, int TWSTEP>
__device__ __forceinline__ void fft_stage(cplx<R>* s, const cplx<R>* tw, int tid) {
    constexpr int NBF = (N / RAD) * NL;
    constexpr int B = (NBF + T - 1) / T;
    constexpr int STRIDE = N / RAD;
    cplx<R> v[B][RAD];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int g = tid + b * T;
        if (NBF % T == 0 || g < NBF) {
            const int l = g % NL, j = g / NL;
#pragma unroll
            for (int m = 0; m < RAD; ++m) v[b][m] = s[(j + m * STRIDE) * ES + l];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int g = tid + b * T;
        if (NBF % T == 0 || g < NBF) {
            const int l = g % NL, j = g / NL;
            const int k = j % NS;
            if (NS > 1) {
#pragma unroll
                for (int m = 1; m < RAD; ++m)
                    v[b][m] = cmul(v[b][m], tw[TWSTEP * m * k * (N / (NS * RAD))]);
            }
            Dft<R, RAD>::run(v[b]);
            const int base = (j / NS) * NS * RAD + k;
#pragma unroll
            for (int m = 0; m < RAD; ++m) s[(base + m * NS) * ES + l] = v[b][m];
        }
    }
    __syncthreads();
}

template <typename R, int N, int NL, int ES, int T, int TWSTEP, int S, int NS>
__device__ __forceinline__ void fft_stages(cplx<R>* s, const cplx<R>* tw, int tid) {
    constexpr Radices P = make_radices(N);
    if constexpr (S < P.n) {
        constexpr int RAD = P.r[S];
        fft_stage<R, N, NL, ES, T, RAD, NS, TWSTEP>(s, tw, tid);
        fft_stages<R, N, NL, ES, T, TWSTEP, S + 1, NS * RAD>(s, tw, tid);
    }
}

// In-place forward FFT of NL lines held in LDS.  The caller must __syncthreads()
// after filling s; on return s holds the result and all threads are synchronised.
template <typename R, int N, int NL, int ES, int T, int TWSTEP>
__device__ __forceinline__ void lds_fft(cplx<R>* s, const cplx<R>* tw, int tid) {
    static_assert(make_radices(N).n > 0, "unsupported FFT length");
    fft_stages<R, N, NL, ES, T, TWSTEP, 0, 1>(s, tw, tid);
}

// ------------------------------------------------------------------ register-fed first / last stage
// Register footprint of the first stage: B0 butterflies of RAD0 points per thread.
template <int N, int NL, int T> struct First {
    static constexpr Radices P = make_radices(N);
    static constexpr int RAD = P.r[0];
    static constexpr int STRIDE = N / RAD;
    static constexpr int NBF = STRIDE * NL;
    static constexpr int B = (NBF + T - 1) / T;
    static constexpr int NST = P.n;
};

// Load the first-stage inputs of the lines held by this thread: butterfly g = tid + b*T covers
// line l = g % NL and elements j + m*STRIDE, j = g / NL.  ld(j, l) returns one element.
template <typename R, int N, int NL, int T, class V, class LD>
__device__ __forceinline__ void first_load(V& v, LD&& ld, int tid) {
    using F = First<N, NL, T>;
#pragma unroll
    for (int b = 0; b < F::B; ++b) {
        const int g = tid + b * T;
        if (F::NBF % T == 0 || g < F::NBF) {
            const int l = g % NL, j = g / NL;
#pragma unroll
            for (int m = 0; m < F::RAD; ++m) v[b][m] = ld(j + m * F::STRIDE, l);
        }
    }
}

// First stage (no twiddles) from registers into LDS.  The caller synchronises before (LDS reuse)
// and this function synchronises after.
template <typename R, int N, int NL, int ES, int T, class V>
__device__ __forceinline__ void first_stage_regs(cplx<R>* s, V& v, int tid) {
    using F = First<N, NL, T>;
#pragma unroll
    for (int b = 0; b < F::B; ++b) {
        const int g = tid + b * T;
        if (F::NBF % T == 0 || g < F::NBF) {
            const int l = g % NL, j = g / NL;
            Dft<R, F::RAD>::run(v[b]);
#pragma unroll
            for (int m = 0; m < F::RAD; ++m) s[(j * F::RAD + m) * ES + l] = v[b][m];
        }
    }
    __syncthreads();
}

// Last stage from LDS straight to the caller's store st(k, l, value) (natural order k).
template <typename R, int N, int NL, int ES, int T, int TWSTEP, class ST>
__device__ __forceinline__ void last_stage_store(const cplx<R>* s, const cplx<R>* tw, ST&& st, int tid) {
    constexpr Radices P = make_radices(N);
    constexpr int RAD = P.r[P.n - 1];
    constexpr int NS = N / RAD;
    constexpr int NBF = NS * NL;
    constexpr int B = (NBF + T - 1) / T;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int g = tid + b * T;
        if (NBF % T == 0 || g < NBF) {
            const int l = g % NL, j = g / NL;  // j < NS: base index k = j
            cplx<R> v[RAD];
#pragma unroll
            for (int m = 0; m < RAD; ++m) v[m] = s[(j + m * NS) * ES + l];
#pragma unroll
            for (int m = 1; m < RAD; ++m) v[m] = cmul(v[m], tw[TWSTEP * m * j * (N / (NS * RAD))]);
            Dft<R, RAD>::run(v);
#pragma unroll
            for (int m = 0; m < RAD; ++m) st(j + m * NS, l, v[m]);
        }
    }
}

// Stages 1 .. n-2 in LDS (stage 0 done by first_stage_regs, stage n-1 by last_stage_store).
template <typename R, int N, int NL, int ES, int T, int TWSTEP, int S, int NS>
__device__ __forceinline__ void mid_stages(cplx<R>* s, const cplx<R>* tw, int tid) {
    constexpr Radices P = make_radices(N);
    if constexpr (S < P.n - 1) {
        constexpr int RAD = P.r[S];
        fft_stage<R, N, NL, ES, T, RAD, NS, TWSTEP>(s, tw, tid);
        mid_stages<R, N, NL, ES, T, TWSTEP, S + 1, NS * RAD>(s, tw, tid);
    }
}

// launch geometry shared by the host launchers and kernels
template <typename R> struct Geo {
    static constexpr int NL = sizeof(R) == 4 ? 16 : 8;  // 128-byte column segments
};
constexpr int clamp_i(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
// threads per block for NL lines of length N: about 16 elements per thread
constexpr int threads_for(int N, int NL) { return clamp_i(N * NL / 16, 64, 256); }

}  // namespace pyl
