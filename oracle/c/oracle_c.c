/* C restatements used as the FP64 oracle at large sizes and as the CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into oracle/liboracle.so and
 * loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 *  - oracle_deposit_f64: the NGP/CIC/TSC/PCS stencils of MAS_library.pyx:123-545
 *    (same weights and index rules as MAS_c.c:8-266) evaluated in float64 with a
 *    float64 grid, floor-based periodic indices.  Serial and deterministic;
 *    oracle_deposit_f64_par gives the same bits on several threads (grid planes split among them).
 *  - oracle_pkloop_f64: the Pk mode loop, Pk_library.pyx:314-381, in float64
 *    (float64 spectrum, float64 window factor).
 *  - oracle_pkloop_f32ref: the same loop as the reference compiles it: complex64
 *    spectrum multiplied in place by a float32 window factor (Pk_library.pyx:275,
 *    354-355), float64 histograms.  Serial, like the reference (the prange import
 *    is commented out at Pk_library.pyx:7).  This is the CPU baseline's mode loop.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline long wrapi(long i, long n) { long r = i % n; return r < 0 ? r + n : r; }

/* 1-D stencil: fills idx[S], w[S] for one axis, returns S. */
static int stencil(double dist, int scheme, int dims, long *idx, double *w)
{
    int j;
    if (scheme == 1) {                       /* NGP, MAS_library.pyx:290-291 */
        idx[0] = wrapi((long)floor(dist + 0.5), dims); w[0] = 1.0; return 1;
    }
    if (scheme == 2) {                       /* CIC, MAS_library.pyx:151-157 */
        double b = floor(dist), u = dist - b;
        idx[0] = wrapi((long)b, dims); idx[1] = wrapi((long)b + 1, dims);
        w[0] = 1.0 - u; w[1] = u; return 2;
    }
    if (scheme == 3) {                       /* TSC, MAS_library.pyx:393-399 */
        double mn = floor(dist - 1.5);
        for (j = 0; j < 3; j++) {
            double c = mn + j + 1, diff = fabs(c - dist);
            idx[j] = wrapi((long)c, dims);
            w[j] = diff < 0.5 ? 0.75 - diff * diff : (diff < 1.5 ? 0.5 * (1.5 - diff) * (1.5 - diff) : 0.0);
        }
        return 3;
    }
    {                                        /* PCS, MAS_library.pyx:486-492 */
        double mn = floor(dist - 2.0);
        for (j = 0; j < 4; j++) {
            double c = mn + j + 1, diff = fabs(c - dist);
            idx[j] = wrapi((long)c, dims);
            w[j] = diff < 1.0 ? (4.0 - 6.0 * diff * diff + 3.0 * diff * diff * diff) / 6.0
                 : (diff < 2.0 ? (2.0 - diff) * (2.0 - diff) * (2.0 - diff) / 6.0 : 0.0);
        }
        return 4;
    }
}

void oracle_deposit_f64(const float *pos, const float *W, long n, int axes, int dims,
                        double BoxSize, int scheme, double multiplicity, double *grid)
{
    double inv = dims / BoxSize;
    long ix[3][4]; double wx[3][4]; int S = 1, a, l, m, q;
    long p;
    for (p = 0; p < n; p++) {
        double ww = (W ? (double)W[p] : 1.0) * multiplicity;
        for (a = 0; a < axes; a++) S = stencil((double)pos[p * axes + a] * inv, scheme, dims, ix[a], wx[a]);
        if (axes == 3) {
            for (l = 0; l < S; l++) for (m = 0; m < S; m++) {
                double wlm = wx[0][l] * wx[1][m] * ww;
                long row = (ix[0][l] * dims + ix[1][m]) * dims;
                for (q = 0; q < S; q++) grid[row + ix[2][q]] += wlm * wx[2][q];
            }
        } else {
            for (l = 0; l < S; l++) for (m = 0; m < S; m++)
                grid[ix[0][l] * dims + ix[1][m]] += wx[0][l] * wx[1][m] * ww;
        }
    }
}

/* The same deposit on `threads` OpenMP threads, 3-D only.  Thread t owns the grid planes
 * ix in [t*dims/threads, (t+1)*dims/threads) and walks every particle in input order, adding only
 * the stencil cells of its planes: every cell receives exactly the serial function's terms in the
 * serial order, so the result is bit-identical to oracle_deposit_f64. */
void oracle_deposit_f64_par(const float *pos, const float *W, long n, int dims, double BoxSize,
                            int scheme, double multiplicity, double *grid, int threads)
{
    double inv = dims / BoxSize;
    if (threads < 1) threads = 1;
    if (threads > dims) threads = dims;
#pragma omp parallel num_threads(threads)
    {
        int t = 0, nt = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        nt = omp_get_num_threads();
#endif
        long x0 = (long)dims * t / nt, x1 = (long)dims * (t + 1) / nt;
        long ix[3][4]; double wx[3][4]; int S = 1, a, l, m, q;
        long p;
        for (p = 0; p < n; p++) {
            double ww;
            S = stencil((double)pos[p * 3] * inv, scheme, dims, ix[0], wx[0]);
            for (l = 0; l < S; l++) if (ix[0][l] >= x0 && ix[0][l] < x1) break;
            if (l == S) continue;
            for (a = 1; a < 3; a++) S = stencil((double)pos[p * 3 + a] * inv, scheme, dims, ix[a], wx[a]);
            ww = (W ? (double)W[p] : 1.0) * multiplicity;
            for (l = 0; l < S; l++) {
                if (ix[0][l] < x0 || ix[0][l] >= x1) continue;
                for (m = 0; m < S; m++) {
                    double wlm = wx[0][l] * wx[1][m] * ww;
                    long row = (ix[0][l] * dims + ix[1][m]) * dims;
                    for (q = 0; q < S; q++) grid[row + ix[2][q]] += wlm * wx[2][q];
                }
            }
        }
    }
}

static inline double mas_corr(double x, int p) { return x == 0.0 ? 1.0 : pow(x / sin(x), p); }
static inline long isqrt_l(long v) {
    long r = (long)sqrt((double)v);
    while (r * r > v) r--;
    while ((r + 1) * (r + 1) <= v) r++;
    return r;
}

/* Shared body of the mode loop, Pk_library.pyx:314-381.  CT = spectrum scalar type,
 * FT = window-factor type.  Histograms are float64 and must be zeroed by the caller. */
#define PKLOOP_BODY(CT, FT)                                                              \
    int middle = dims / 2, kxx, kyy, kzz, even = (dims % 2 == 0);                       \
    int kmax_par = middle;                                                              \
    long nz = middle + 1;                                                               \
    double prefact = M_PI / dims, c0, c1, c2;                                           \
    for (kxx = 0; kxx < dims; kxx++) {                                                  \
        int kx = kxx > middle ? kxx - dims : kxx;                                       \
        c0 = mas_corr(prefact * kx, p);                                                 \
        for (kyy = 0; kyy < dims; kyy++) {                                              \
            int ky = kyy > middle ? kyy - dims : kyy;                                   \
            c1 = mas_corr(prefact * ky, p);                                             \
            for (kzz = 0; kzz <= middle; kzz++) {                                       \
                int kz = kzz, k_par, k_per, k_index, index_2D;                          \
                double k, mu, mu2, re, im, d2, ph;                                      \
                FT fac;                                                                 \
                long o;                                                                 \
                c2 = mas_corr(prefact * kz, p);                                         \
                if (kz == 0 || (kz == middle && even)) {                                \
                    if (kx < 0) continue;                                               \
                    else if (kx == 0 || (kx == middle && even)) { if (ky < 0) continue; } \
                }                                                                       \
                k = sqrt((double)(kx * kx + ky * ky + kz * kz));                        \
                k_index = (int)isqrt_l((long)kx * kx + (long)ky * ky + (long)kz * kz);  \
                if (axis == 0) { k_par = kx; k_per = (int)isqrt_l((long)ky * ky + (long)kz * kz); } \
                else if (axis == 1) { k_par = ky; k_per = (int)isqrt_l((long)kx * kx + (long)kz * kz); } \
                else { k_par = kz; k_per = (int)isqrt_l((long)kx * kx + (long)ky * ky); } \
                mu = (k == 0) ? 0.0 : k_par / k;                                        \
                mu2 = mu * mu;                                                          \
                if (k_par < 0) k_par = -k_par;                                          \
                fac = (FT)(c0 * c1 * c2);                                               \
                o = ((long)kxx * dims + kyy) * nz + kzz;                                \
                {                                                                       \
                    CT r = dk[2 * o] * fac, i_ = dk[2 * o + 1] * fac;                    \
                    re = r; im = i_;                                                    \
                }                                                                       \
                d2 = re * re + im * im;                                                 \
                ph = atan2(re, sqrt(d2));                                               \
                if (k <= middle) { k1D[k_par] += k_par; Pk1D[k_par] += d2; N1D[k_par] += 1.0; } \
                index_2D = (kmax_par + 1) * k_per + k_par;                              \
                Pk2D[index_2D] += d2; N2D[index_2D] += 1.0;                             \
                k3D[k_index] += k;                                                      \
                P3D[3 * k_index + 0] += d2;                                             \
                P3D[3 * k_index + 1] += d2 * (3.0 * mu2 - 1.0) / 2.0;                   \
                P3D[3 * k_index + 2] += d2 * (35.0 * mu2 * mu2 - 30.0 * mu2 + 3.0) / 8.0; \
                phase[k_index] += ph * ph;                                              \
                N3D[k_index] += 1.0;                                                    \
            }                                                                           \
        }                                                                               \
    }

void oracle_pkloop_f64(const double *dk, int dims, int axis, int p,
                       double *k1D, double *Pk1D, double *N1D, double *Pk2D, double *N2D,
                       double *k3D, double *P3D, double *phase, double *N3D)
{
    PKLOOP_BODY(double, double)
}

void oracle_pkloop_f32ref(const float *dk, int dims, int axis, int p,
                          double *k1D, double *Pk1D, double *N1D, double *Pk2D, double *N2D,
                          double *k3D, double *P3D, double *phase, double *N3D)
{
    PKLOOP_BODY(float, float)
}

/* Linear-response variance of the mode-loop sums under white noise on the raw spectrum.
 * A float32 pipeline returns delta_k + e_k with e_k white (E|e_k|^2 = s2).  To first order
 * |C (delta_k + e_k)|^2 - |C delta_k|^2 = 2 C^2 Re(conj(delta_k) e_k), whose variance is
 * 2 C^2 |C delta_k|^2 s2 (C: the window correction of the mode).  Per bin this accumulates
 * sum 2 C^2 |C delta_k|^2 L_l(mu)^2 over the same independent modes and bins as the loop above
 * (Pk_library.pyx:314-381): V1D (1-D), V2D (2-D), V3D (kmax+1, 3) for l = 0, 2, 4; the caller
 * multiplies by s2.  The phase term ph = atan(re / r), r = |C delta|, moves by
 * im (im dre - re dim) / (r (r^2 + re^2)) with (dre, dim) = C e, so ph^2 has variance
 * 2 ph^2 C^2 s2 im^2 / (r^2 + re^2)^2: Vph (kmax+1).  Planes of kx are split among `threads` threads, each with its own
 * histograms, summed at the end. */
void oracle_pkloop_var(const double *dk, int dims, int axis, int p,
                       double *V1D, double *V2D, double *V3D, double *Vph, int threads)
{
    int middle = dims / 2, even = (dims % 2 == 0), kmax_par = middle;
    long nz = middle + 1;
    long n1 = kmax_par + 1, n2 = (long)(kmax_par + 1) * (isqrt_l(2L * middle * middle) + 1);
    long n3 = isqrt_l(3L * middle * middle) + 1;
    double prefact = M_PI / dims;
    int T = threads > 0 ? threads : 1, t;
    size_t per = (size_t)(n1 + n2 + 4 * n3);
    double *buf = (double *)calloc((size_t)T * per, sizeof(double));
    if (!buf) return;
#pragma omp parallel for num_threads(T) schedule(static)
    for (t = 0; t < T; t++) {
        double *b1 = buf + (size_t)t * per, *b2 = b1 + n1, *b3 = b2 + n2, *bp = b3 + 3 * n3;
        int kxx, kyy, kzz;
        for (kxx = t; kxx < dims; kxx += T) {
            int kx = kxx > middle ? kxx - dims : kxx;
            double c0 = mas_corr(prefact * kx, p);
            for (kyy = 0; kyy < dims; kyy++) {
                int ky = kyy > middle ? kyy - dims : kyy;
                double c1 = mas_corr(prefact * ky, p);
                for (kzz = 0; kzz <= middle; kzz++) {
                    int kz = kzz, k_par, k_per, k_index;
                    double k, mu2, C, re, im, v, l2, l4, d2, ph;
                    long o;
                    if (kz == 0 || (kz == middle && even)) {
                        if (kx < 0) continue;
                        else if (kx == 0 || (kx == middle && even)) { if (ky < 0) continue; }
                    }
                    k = sqrt((double)(kx * kx + ky * ky + kz * kz));
                    k_index = (int)isqrt_l((long)kx * kx + (long)ky * ky + (long)kz * kz);
                    if (axis == 0) { k_par = kx; k_per = (int)isqrt_l((long)ky * ky + (long)kz * kz); }
                    else if (axis == 1) { k_par = ky; k_per = (int)isqrt_l((long)kx * kx + (long)kz * kz); }
                    else { k_par = kz; k_per = (int)isqrt_l((long)kx * kx + (long)ky * ky); }
                    mu2 = (k == 0) ? 0.0 : (k_par / k) * (k_par / k);
                    if (k_par < 0) k_par = -k_par;
                    C = c0 * c1 * mas_corr(prefact * kz, p);
                    o = ((long)kxx * dims + kyy) * nz + kzz;
                    re = dk[2 * o] * C; im = dk[2 * o + 1] * C;
                    d2 = re * re + im * im;
                    v = 2.0 * C * C * d2;
                    if (k <= middle) b1[k_par] += v;
                    b2[(long)(kmax_par + 1) * k_per + k_par] += v;
                    l2 = (3.0 * mu2 - 1.0) / 2.0;
                    l4 = (35.0 * mu2 * mu2 - 30.0 * mu2 + 3.0) / 8.0;
                    b3[3 * k_index + 0] += v;
                    b3[3 * k_index + 1] += v * l2 * l2;
                    b3[3 * k_index + 2] += v * l4 * l4;
                    ph = atan2(re, sqrt(d2));
                    if (d2 > 0.0) bp[k_index] += 2.0 * ph * ph * C * C * im * im / ((d2 + re * re) * (d2 + re * re));
                }
            }
        }
    }
    for (t = 0; t < T; t++) {
        const double *b1 = buf + (size_t)t * per, *b2 = b1 + n1, *b3 = b2 + n2, *bp = b3 + 3 * n3;
        long i;
        for (i = 0; i < n1; i++) V1D[i] += b1[i];
        for (i = 0; i < n2; i++) V2D[i] += b2[i];
        for (i = 0; i < 3 * n3; i++) V3D[i] += b3[i];
        for (i = 0; i < n3; i++) Vph[i] += bp[i];
    }
    free(buf);
}
