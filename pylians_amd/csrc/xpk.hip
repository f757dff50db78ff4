// Auto and cross power spectra of several fields (Pk_library.XPk, Pk_library.pyx:534-798) on gfx950.
//
// Every field is transformed (r2c, stored half spectrum); one binning kernel then walks the
// independent modes once per group of (field a, field b) pairs, applies each field's window
// (its own MAS, Pk_library.pyx:688-697) and accumulates Re(d_a conj(d_b)) into the 3-D
// multipoles, 1-D and 2-D histograms (Pk_library.pyx:699-728) in float64.  The block/column
// assignment is the one of the fused Pk pass: lines run along y (line of sight x) or x, and each
// persistent block owns one line coordinate u (k_par = |k_u|) or one kz block (line of sight z),
// so its 2-D window stays in LDS.
#include <math.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/pylians_hip.h"
#include "fft.hpp"
#include "pk.hpp"
#include "util.hpp"

namespace pyl {

const double* pk_window_table(int N, int p);

constexpr int XMAXF = 8;   // fields per call
constexpr int XMAXC = 4;   // pairs per binning launch

struct XBinArgs {
    const void* spec[XMAXF];
    const double* win[XMAXF];
    int ca[XMAXC], cb[XMAXC];
    int nc;
    int nf, fl[XMAXF];       // distinct fields of the launch's pairs
    int pa[XMAXC], pb[XMAXC];  // pair c = (fl[pa[c]], fl[pb[c]])
    long long P;
    int N, axis, last, fixk, nchunk, kmax_par, kmax_per, n1, n3, n2;
    double* s3;  // [nc][n3][3]
    double* s1;  // [nc][n1]
    double* s2;  // [nc][n2]
    // modes binned per 3-D / 1-D / 2-D bin (Pk_library.pyx:713-728), counted by the first launch
    // only (every launch walks the same modes); nullptr in the others
    double* gc3;  // [n3]
    double* gc1;  // [n1]
    double* gc2;  // [n2]
};

// NF: register slots for the distinct fields of one launch (>= their number); NC: its pairs
// 1 / (mode count) of every 2-D bin of a geometry (inf for an empty bin, whose sum is 0: NaN, as
// 0/0), cached per (grid size, line of sight), which fix the counts
static const std::vector<double>& inv_counts(const PkGeometry* g) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::vector<double>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto& v = cache[std::make_pair(g->d.N, g->axis)];
    if (v.size() != g->h_n2.size()) {
        v.resize(g->h_n2.size());
        for (size_t i = 0; i < v.size(); ++i) v[i] = 1.0 / (double)g->h_n2[i];
    }
    return v;
}

template <typename R, int NL, int T, int NF, int NC>
__global__ __launch_bounds__(T) void k_xbin(XBinArgs a) {
    constexpr int TPC = T / NL;
    extern __shared__ double hs[];
    const int N = a.N, nc = a.nc, nper = a.kmax_per + 1;
    const int wlen = (a.fixk ? NL : 1) * nper;
    double* h3 = hs;                       // [n3][nc][3]
    double* h1 = h3 + a.n3 * nc * 3;       // [nc][n1]
    double* h2 = h1 + nc * a.n1;           // [nc][wlen]
    const int nsum = a.n3 * nc * 3 + nc * a.n1 + nc * wlen;
    unsigned* q3 = (unsigned*)(hs + nsum);   // mode counts [n3] | [n1] | [wlen]
    unsigned* q1 = q3 + a.n3;
    unsigned* q2 = q1 + a.n1;
    const bool cnt = a.gc3 != nullptr;
    const int tid = threadIdx.x;
    for (int i = tid; i < nsum; i += T) hs[i] = 0.0;
    for (int i = tid; i < a.n3 + a.n1 + wlen; i += T) q3[i] = 0u;
    __syncthreads();
    const int middle = N / 2;
    const bool even = (N % 2) == 0;
    const long long P = a.P;
    const int nkzb = (int)((P + NL - 1) / NL);
    const long long EST = a.last == 0 ? (long long)N * P : P;
    const long long UST = a.last == 0 ? P : (long long)N * P;
    const int l = tid % NL, jrow = tid / NL;
    const int E = (N + TPC - 1) / TPC;
    int u0, u1, kzb0, kzb1;
    if (a.fixk) {
        const int kzb = blockIdx.x / a.nchunk, ch = blockIdx.x % a.nchunk;
        const int U = (N + a.nchunk - 1) / a.nchunk;
        u0 = ch * U;
        u1 = min(N, u0 + U);
        kzb0 = kzb;
        kzb1 = kzb + 1;
    } else {
        u0 = blockIdx.x;
        u1 = u0 + 1;
        kzb0 = 0;
        kzb1 = nkzb;
    }
    const int kpar_u = a.fixk ? 0 : abs(kfreq(u0, N));
    // distinct fields of this launch's pairs (from the host): each is loaded once per mode
    const int nf = a.nf;
    const cplx<R>* sp[NF];
    const double* wp[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const int ff = f < nf ? a.fl[f] : a.fl[0];
        sp[f] = (const cplx<R>*)a.spec[ff];
        wp[f] = a.win[ff];
    }
    for (int u = u0; u < u1; ++u)
        for (int kzb = kzb0; kzb < kzb1; ++kzb) {
            const int kz = kzb * NL + l;
            if (kz >= P) continue;
            const long long base = (long long)u * UST + kz;
            const int ku = kfreq(u, N);
            const bool zplane = (kz == 0) || (kz == middle && even);
            // product of the window factors of the two fixed coordinates (u and kz), per field; a mode's
            // window is this times the factor of its k_j (Pk_library.pyx:688-697, in float64, rounded
            // once to the working precision)
            double wuz[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f)
                if (f < nf) wuz[f] = wp[f][u] * wp[f][kz];
            // |k|^2, k_perp^2 of the thread's first mode in the column and their bins floor(sqrt); along
            // the run |k_j| changes by one per step (across j = N/2 too), so each bin moves by at most one
            int bin = 0, wb = 0;
            {
                const int j0 = jrow * E;
                if (j0 < N) {
                    const int kj = kfreq(j0, N);
                    const int kx = a.last == 0 ? kj : ku, ky = a.last == 0 ? ku : kj;
                    bin = isqrt24(kx * kx + ky * ky + kz * kz);
                    wb = isqrt24(a.axis == 0 ? ky * ky + kz * kz : (a.axis == 1 ? kx * kx + kz * kz : kx * kx + ky * ky));
                }
            }
            int cur3 = -1, cur2 = -1;
            unsigned m3 = 0u, m1 = 0u, m2 = 0u;   // modes of the current runs
            double acc[NC][3], b1[NC], b2[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                acc[c][0] = acc[c][1] = acc[c][2] = 0.0;
                b1[c] = b2[c] = 0.0;
            }
            // QB modes per batch: every field value and window entry of the batch is loaded before
            // any of them is used
            constexpr int QB = 4;
            for (int q0 = 0; q0 < E; q0 += QB) {
                cplx<R> v[QB][NF];
                double wj[QB][NF];
#pragma unroll
                for (int b = 0; b < QB; ++b) {
                    const int j = jrow * E + q0 + b;
                    const bool in = q0 + b < E && j < N;
#pragma unroll
                    for (int f = 0; f < NF; ++f)
                        if (f < nf) {
                            v[b][f] = in ? sp[f][base + j * EST] : mk<R>(0, 0);
                            wj[b][f] = in ? wp[f][j] : 0.0;
                        }
                }
#pragma unroll
                for (int b = 0; b < QB; ++b) {
                    const int j = jrow * E + q0 + b;
                    if (q0 + b >= E || j >= N) break;
                    const int kj = kfreq(j, N);
                    const int kx = a.last == 0 ? kj : ku;
                    const int ky = a.last == 0 ? ku : kj;
                    const int k2 = kx * kx + ky * ky + kz * kz;
                    const int kp2 = a.axis == 0 ? ky * ky + kz * kz : (a.axis == 1 ? kx * kx + kz * kz : kx * kx + ky * ky);
                    bin += __mul24(bin + 1, bin + 1) <= k2 ? 1 : 0;
                    bin -= __mul24(bin, bin) > k2 ? 1 : 0;
                    wb += __mul24(wb + 1, wb + 1) <= kp2 ? 1 : 0;
                    wb -= __mul24(wb, wb) > kp2 ? 1 : 0;
                    if (zplane && (kx < 0 || ((kx == 0 || (kx == middle && even)) && ky < 0))) continue;
                    const int kps = a.axis == 0 ? kx : (a.axis == 1 ? ky : kz);
                    const R mu2 = k2 == 0 ? (R)0 : (R)(kps * kps) * recip((R)k2);
                    const R L2 = ((R)3 * mu2 - (R)1) * (R)0.5;
                    const R L4 = (((R)35 * mu2 - (R)30) * mu2 + (R)3) * (R)0.125;
                    // each field's value times its window W(x) W(y) W(z)
                    cplx<R> w[NF];
#pragma unroll
                    for (int f = 0; f < NF; ++f)
                        if (f < nf) w[f] = scale(v[b][f], (R)(wj[b][f] * wuz[f]));
                    R val[NC];
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        cplx<R> va = w[0], vb = w[0];
#pragma unroll
                        for (int f = 1; f < NF; ++f) {
                            if (a.pa[c] == f) va = w[f];
                            if (a.pb[c] == f) vb = w[f];
                        }
                        val[c] = va.x * vb.x + va.y * vb.y;
                    }
                    if (bin != cur3) {
                        if (cur3 >= 0) {
#pragma unroll
                            for (int c = 0; c < NC; ++c)
                                for (int m = 0; m < 3; ++m) atomicAdd(&h3[(cur3 * nc + c) * 3 + m], acc[c][m]);
                            if (cnt) atomicAdd(&q3[cur3], m3);
                        }
                        cur3 = bin;
                        m3 = 0u;
#pragma unroll
                        for (int c = 0; c < NC; ++c) acc[c][0] = acc[c][1] = acc[c][2] = 0.0;
                    }
                    ++m3;
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        acc[c][0] += (double)val[c];
                        acc[c][1] += (double)(val[c] * L2);
                        acc[c][2] += (double)(val[c] * L4);
                    }
                    if (k2 <= middle * middle) {
                        ++m1;
#pragma unroll
                        for (int c = 0; c < NC; ++c) b1[c] += (double)val[c];
                    }
                    const int kper = wb;
                    const int w2 = (a.fixk ? l * nper : 0) + kper;
                    if (w2 != cur2) {
                        if (cur2 >= 0) {
#pragma unroll
                            for (int c = 0; c < NC; ++c) atomicAdd(&h2[c * wlen + cur2], b2[c]);
                            if (cnt) atomicAdd(&q2[cur2], m2);
                        }
                        cur2 = w2;
                        m2 = 0u;
#pragma unroll
                        for (int c = 0; c < NC; ++c) b2[c] = 0.0;
                    }
                    ++m2;
#pragma unroll
                    for (int c = 0; c < NC; ++c) b2[c] += (double)val[c];
                }
            }
            const int kpc = a.fixk ? kz : kpar_u;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                if (cur3 >= 0)
                    for (int m = 0; m < 3; ++m) atomicAdd(&h3[(cur3 * nc + c) * 3 + m], acc[c][m]);
                if (cur2 >= 0) atomicAdd(&h2[c * wlen + cur2], b2[c]);
                if (m1 != 0u) atomicAdd(&h1[c * a.n1 + kpc], b1[c]);
            }
            if (cnt) {
                if (cur3 >= 0) atomicAdd(&q3[cur3], m3);
                if (cur2 >= 0) atomicAdd(&q2[cur2], m2);
                if (m1 != 0u) atomicAdd(&q1[kpc], m1);
            }
        }
    __syncthreads();
    if (cnt) {
        for (int i = tid; i < a.n3; i += T)
            if (q3[i]) atomicAdd(&a.gc3[i], (double)q3[i]);
        for (int i = tid; i < a.n1; i += T)
            if (q1[i]) atomicAdd(&a.gc1[i], (double)q1[i]);
        for (int i = tid; i < wlen; i += T) {
            if (!q2[i]) continue;
            const int kper = i % nper;
            const int kpar = a.fixk ? kzb0 * NL + i / nper : kpar_u;
            if (kpar <= a.kmax_par) atomicAdd(&a.gc2[(long long)(a.kmax_par + 1) * kper + kpar], (double)q2[i]);
        }
    }
    for (int i = tid; i < a.n3 * nc * 3; i += T) {
        if (h3[i] == 0.0) continue;
        const int bin = i / (nc * 3), c = (i / 3) % nc, m = i % 3;
        atomicAdd(&a.s3[((long long)c * a.n3 + bin) * 3 + m], h3[i]);
    }
    for (int i = tid; i < nc * a.n1; i += T)
        if (h1[i] != 0.0) atomicAdd(&a.s1[i], h1[i]);
    for (int i = tid; i < nc * wlen; i += T) {
        if (h2[i] == 0.0) continue;
        const int c = i / wlen, r = i % wlen, kper = r % nper;
        const int kpar = a.fixk ? kzb0 * NL + r / nper : kpar_u;
        if (kpar <= a.kmax_par) atomicAdd(&a.s2[(long long)c * a.n2 + (long long)(a.kmax_par + 1) * kper + kpar], h2[i]);
    }
}

template <typename R, int NL, int T, int NF>
static void launch_xbin_nc(const XBinArgs& a, int grid, size_t shm, hipStream_t st) {
    switch (a.nc) {
        case 1: hipLaunchKernelGGL((k_xbin<R, NL, T, NF, 1>), dim3((unsigned)grid), dim3(T), shm, st, a); break;
        case 2: hipLaunchKernelGGL((k_xbin<R, NL, T, NF, 2>), dim3((unsigned)grid), dim3(T), shm, st, a); break;
        case 3: hipLaunchKernelGGL((k_xbin<R, NL, T, NF, 3>), dim3((unsigned)grid), dim3(T), shm, st, a); break;
        default: hipLaunchKernelGGL((k_xbin<R, NL, T, NF, 4>), dim3((unsigned)grid), dim3(T), shm, st, a); break;
    }
}

// With the line of sight along z (NL = 8) the blocks hold a 2-D histogram row per column and pair,
// which leaves LDS for one block per CU: up to 4 distinct fields they run 2T threads (two fields of
// 512^3: 3.13 -> 2.78 ms per call, four: 9.1 -> 7.9), while 8 fields keep T, where 2T threads spill.
template <typename R, int NL, int T>
static int launch_xbin(const XBinArgs& a, int grid, size_t shm, hipStream_t st) {
    PYL_CHECK(a.nc >= 1 && a.nc <= XMAXC && a.nf >= 1 && a.nf <= XMAXF, E_ARG, "bad pair group");
    constexpr int TW = NL == 8 ? 2 * T : T;
    if (a.nf <= 2)
        launch_xbin_nc<R, NL, TW, 2>(a, grid, shm, st);
    else if (a.nf <= 4)
        launch_xbin_nc<R, NL, TW, 4>(a, grid, shm, st);
    else
        launch_xbin_nc<R, NL, T, 8>(a, grid, shm, st);
    PYL_HIP(hipGetLastError());
    return OK;
}

static int num_cus_x() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
        n = 256;
    return n;
}

template <typename R, int NL, int T>
static int xpk_run_k(const void* const* deltas, int F, int N, double BoxSize, int axis, const int* mas_p, double* out,
                     hipStream_t st) {
    int rc = OK;
    const int64_t P = half_len(N);
    const PkGeometry* g = pk_geometry(N, axis, &rc);
    if (rc) return rc;
    const PkDims& d = g->d;
    const size_t nspec = (size_t)N * N * P;
    cplx<R>* spec = (cplx<R>*)workspace(WS_SPEC, sizeof(cplx<R>) * nspec * F, &rc);
    if (rc) return rc;
    for (int f = 0; f < F; ++f) {
        rc = fft3d_r2c<R>((const R*)deltas[f], spec + nspec * f, N, P, st);
        if (rc) return rc;
    }
    std::vector<int> pa, pb;
    for (int i = 0; i < F; ++i) {
        pa.push_back(i);
        pb.push_back(i);
    }
    for (int i = 0; i < F; ++i)
        for (int j = i + 1; j < F; ++j) {
            pa.push_back(i);
            pb.push_back(j);
        }
    const int C = (int)pa.size();
    const int per = 3 * d.n3() + d.n1() + d.n2();
    // pair sums, then the mode counts n3 | n1 | n2
    const size_t nsums = (size_t)per * C + d.n3() + d.n1() + d.n2();
    double* sums = (double*)workspace(WS_SUMS, sizeof(double) * nsums, &rc);
    if (rc) return rc;
    rc = zero_fill(sums, sizeof(double) * nsums, st);
    if (rc) return rc;
    double* gcount = sums + (size_t)per * C;
    XBinArgs a{};
    for (int f = 0; f < F; ++f) {
        a.spec[f] = spec + nspec * f;
        a.win[f] = pk_window_table(N, mas_p[f]);
        PYL_CHECK(a.win[f], E_ALLOC, "window table allocation failed");
    }
    a.P = P;
    a.N = N;
    a.axis = axis;
    a.last = pk_last_axis(axis);
    a.fixk = axis == 2 ? 1 : 0;
    a.kmax_par = d.kmax_par;
    a.kmax_per = d.kmax_per;
    a.n1 = d.n1();
    a.n2 = d.n2();
    a.n3 = d.n3();
    const int nper = d.kmax_per + 1;
    const int wlen = (a.fixk ? NL : 1) * nper;
    // pairs per launch: as many (<= 4) as fit the LDS histograms
    int cmax = XMAXC;
    while (cmax > 1 && (size_t)cmax * (3 * d.n3() + d.n1() + wlen) * 8 + (size_t)(d.n3() + d.n1() + wlen) * 4 >
                           120 * 1024)
        --cmax;
    const int nkzb = (int)((P + NL - 1) / NL);
    a.nchunk = 1;
    int grid = N;
    if (a.fixk) {
        a.nchunk = (num_cus_x() * 4 + nkzb - 1) / nkzb;
        if (a.nchunk > N) a.nchunk = N;
        grid = nkzb * a.nchunk;
    }
    for (int c0 = 0; c0 < C; c0 += cmax) {
        a.nc = C - c0 < cmax ? C - c0 : cmax;
        for (int c = 0; c < a.nc; ++c) {
            a.ca[c] = pa[c0 + c];
            a.cb[c] = pb[c0 + c];
        }
        // a launch group of nc pairs owns nc * per doubles, kind-major: [nc][n3][3] | [nc][n1] | [nc][n2]
        double* base = sums + (size_t)per * c0;
        a.s3 = base;
        a.s1 = base + (size_t)3 * d.n3() * a.nc;
        a.s2 = a.s1 + (size_t)d.n1() * a.nc;
        const size_t shm = sizeof(double) * (size_t)a.nc * (3 * d.n3() + d.n1() + wlen) +
                           sizeof(unsigned) * (size_t)(d.n3() + d.n1() + wlen);
        a.gc3 = c0 == 0 ? gcount : nullptr;
        a.gc1 = c0 == 0 ? gcount + d.n3() : nullptr;
        a.gc2 = c0 == 0 ? gcount + d.n3() + d.n1() : nullptr;
        std::vector<int> fs;   // distinct fields of the launch
        for (int c = 0; c < a.nc; ++c) {
            int ip[2];
            for (int e = 0; e < 2; ++e) {
                const int f = e == 0 ? a.ca[c] : a.cb[c];
                auto it = std::find(fs.begin(), fs.end(), f);
                ip[e] = (int)(it - fs.begin());
                if (it == fs.end()) fs.push_back(f);
            }
            a.pa[c] = ip[0];
            a.pb[c] = ip[1];
        }
        a.nf = (int)fs.size();
        for (int f = 0; f < a.nf; ++f) a.fl[f] = fs[f];
        rc = launch_xbin<R, NL, T>(a, grid, shm, st);
        if (rc) return rc;
    }
    // one copy into the library's pinned buffer (no pageable staging, no per-call host allocation)
    double* hp = (double*)host_pinned(sizeof(double) * nsums, &rc);
    if (rc) return rc;
    PYL_HIP(hipMemcpyAsync(hp, sums, sizeof(double) * nsums, hipMemcpyDeviceToHost, st));
    PYL_HIP(hipStreamSynchronize(st));
    const double* h = hp;
    // Nmodes from the kernel's counts, checked bin by bin against the independent-mode geometry
    // (check_number_modes, Pk_library.pyx:90-102, 785)
    const double* k3c = h + (size_t)per * C;
    const double* k1c = k3c + d.n3();
    const double* k2c = k1c + d.n1();
    {
        int bad = 0;
        double tot = 0.0;
        for (int i = 0; i < d.n3(); ++i) {
            bad += k3c[i] != (double)g->h_n3[i];
            tot += k3c[i];
        }
        for (int i = 0; i < d.n1(); ++i) bad += k1c[i] != (double)g->h_n1[i];
        for (int i = 0; i < d.n2(); ++i) bad += k2c[i] != (double)g->h_n2[i];
        const long long NN = N, own = (N % 2 == 1) ? 1 : 8;
        PYL_CHECK(bad == 0, E_MODES,
                  "Not all modes counted: counted %.0f independent modes, expected %lld (%d bins differ from the "
                  "mode geometry)", tot, (NN * NN * NN - own) / 2 + own, bad);
    }
    // finalize (Pk_library.pyx:730-785)
    const double kF = 2.0 * M_PI / BoxSize, kN = d.middle * kF;
    const double u1 = BoxSize / ((double)N * N);
    const double fact = u1 * u1 * u1;
    const int kmax = d.kmax, kp = d.kmax_par, n2 = d.n2();
    double* o = out;
    for (int i = 0; i < kmax; ++i) o[i] = g->h_k3sum[i + 1] / k3c[i + 1] * kF;
    o += kmax;
    for (int i = 0; i < kmax; ++i) o[i] = k3c[i + 1];
    o += kmax;
    for (int i = 0; i < kp; ++i) o[i] = ((double)(i + 1) * k1c[i + 1] / k1c[i + 1]) * kF;
    double* k1D = o;
    o += kp;
    for (int i = 0; i < kp; ++i) o[i] = k1c[i + 1];
    o += kp;
    // 2-D bin i = (kmax_par + 1) * kper + kpar
    for (int i = 0, r = 0; i < n2; ++r)
        for (int q = 0; q <= kp; ++q, ++i) {
            o[i] = 0.5 * (2 * q + 1) * kF;
            o[n2 + i] = 0.5 * (2 * r + 1) * kF;
        }
    o += 2 * (size_t)n2;
    for (int i = 0; i < n2; ++i) o[i] = k2c[i];
    o += n2;
    // 1/Nmodes2D of the geometry (equal to the kernel's counts, checked above), once per geometry:
    // one division per 2-D bin instead of one per bin and pair
    const std::vector<double>& inv2 = inv_counts(g);
    // the pairs interleaved, as the reference's arrays Pk[kmax][3][F] / Pk1D[kmax_par][F] / Pk2D[n2][F]
    // hold them: P[kmax][3][C], then P1D[kmax_par][C], then P2D[n2][C]
    double* o3 = o;
    double* o1 = o3 + (size_t)3 * kmax * C;
    double* o2 = o1 + (size_t)kp * C;
    for (int c = 0; c < C; ++c) {
        // locate this pair's sums inside its launch group
        const int c0 = (c / cmax) * cmax, nc = (C - c0 < cmax ? C - c0 : cmax), cc = c - c0;
        const double* gb = h + (size_t)per * c0;
        const double* s3 = gb + ((size_t)cc * d.n3()) * 3;
        const double* s1 = gb + (size_t)3 * d.n3() * nc + (size_t)cc * d.n1();
        const double* s2 = gb + (size_t)3 * d.n3() * nc + (size_t)d.n1() * nc + (size_t)cc * n2;
        for (int i = 0; i < kmax; ++i) {
            const double n = k3c[i + 1];
            o3[(size_t)(3 * i + 0) * C + c] = (s3[3 * (i + 1) + 0] / n) * fact;
            o3[(size_t)(3 * i + 1) * C + c] = (s3[3 * (i + 1) + 1] * 5.0 / n) * fact;
            o3[(size_t)(3 * i + 2) * C + c] = (s3[3 * (i + 1) + 2] * 9.0 / n) * fact;
        }
        for (int i = 0; i < kp; ++i) {
            const double n = k1c[i + 1];
            const double kmaxper = sqrt(kN * kN - k1D[i] * k1D[i]);
            o1[(size_t)i * C + c] = s1[i + 1] * fact * (M_PI * kmaxper * kmaxper / n) / ((2.0 * M_PI) * (2.0 * M_PI));
        }
        for (int i = 0; i < n2; ++i) o2[(size_t)i * C + c] = s2[i] * fact * inv2[i];
    }
    return OK;
}


// 16 columns per block (whole 128-byte lines per row read); with the line of sight along z the blocks
// keep a 2-D histogram row per column, and 8 columns let all pairs of two fields share one launch
// (16 needed two, each reading both spectra: 5.3 -> 3.7 ms for two 512^3 fields)
template <typename R>
static int xpk_run(const void* const* deltas, int F, int N, double BoxSize, int axis, const int* mas_p, double* out,
                   hipStream_t st) {
    if (axis == 2) return xpk_run_k<R, 8, 256>(deltas, F, N, BoxSize, axis, mas_p, out, st);
    return xpk_run_k<R, 16, 256>(deltas, F, N, BoxSize, axis, mas_p, out, st);
}

}  // namespace pyl

using namespace pyl;

extern "C" int pyl_xpk_layout(int dims, int nfields, int64_t* out_len) {
    PYL_CHECK(dims > 1 && nfields >= 1 && nfields <= XMAXF, E_ARG, "bad dims / nfields (1..%d)", XMAXF);
    const PkDims d = pk_dims(dims, half_len(dims));
    const int64_t C = (int64_t)nfields * (nfields + 1) / 2;
    if (out_len)
        *out_len = 2 * (int64_t)d.kmax + 2 * (int64_t)d.kmax_par + 3 * (int64_t)d.n2() +
                   C * (3 * (int64_t)d.kmax + d.kmax_par + d.n2());
    return OK;
}

extern "C" int pyl_xpk(const void* const* deltas, int nfields, int dims, double BoxSize, int axis, const int* mas_p,
                       int prec, double* out, void* stream) {
    PYL_CHECK(prec == PYL_F32 || prec == PYL_F64, E_ARG, "prec must be 4 or 8");
    PYL_CHECK(fft_size_supported(dims), E_UNSUPPORTED, "grid size %d not supported", dims);
    PYL_CHECK(nfields >= 1 && nfields <= XMAXF, E_ARG, "nfields must be in [1, %d]", XMAXF);
    PYL_CHECK(axis >= 0 && axis <= 2 && BoxSize > 0 && deltas && mas_p && out, E_ARG, "bad arguments");
    for (int f = 0; f < nfields; ++f) {
        PYL_CHECK(deltas[f], E_ARG, "null field %d", f);
        PYL_CHECK(mas_p[f] >= 0 && mas_p[f] <= 4, E_ARG, "window exponent must be in [0, 4]");
    }
    hipStream_t st = (hipStream_t)stream;
    if (prec == PYL_F32) return xpk_run<float>(deltas, nfields, dims, BoxSize, axis, mas_p, out, st);
    return xpk_run<double>(deltas, nfields, dims, BoxSize, axis, mas_p, out, st);
}
