/* libpylians_hip: C ABI of the MI355X (gfx950) density-assignment -> P(k)/B(k) path.
 *
 * Conventions
 *  - Every function returns 0 on success and a nonzero PYL_E_* code otherwise; the
 *    message is available from pyl_last_error() (thread-local).
 *  - Unless a name ends in a MAS_c.h drop-in (pyl_mas_NGP/CIC/TSC/PCS), pointers are DEVICE
 *    pointers, the caller owns every input and output buffer, and the work is queued on
 *    `stream` (a hipStream_t; NULL = the default stream) without host synchronisation.
 *    The library owns FFT twiddle tables, window tables, bin-geometry tables and grow-only
 *    workspaces, cached per device.
 *  - Calls that share the library workspaces must not run concurrently on different streams
 *    of one device.
 *  - prec selects the working precision: PYL_F32 (float32 grid and FFT, float64 shell sums)
 *    or PYL_F64 (float64 grid, FFT and sums).
 */
#ifndef PYLIANS_HIP_H
#define PYLIANS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYL_ABI_VERSION 1

#define PYL_OK 0
#define PYL_E_ARG 1
#define PYL_E_UNSUPPORTED 2
#define PYL_E_MODES 3
#define PYL_E_ALLOC 4
#define PYL_E_LAUNCH 5

/* mass-assignment schemes; also the window exponent p of Pk_library.MAS_function */
#define PYL_NGP 1
#define PYL_CIC 2
#define PYL_TSC 3
#define PYL_PCS 4

#define PYL_F32 4
#define PYL_F64 8

/* deposit flags */
#define PYL_GRID_F64 1   /* grid is float64 (CICW_d / NGPW_d) */
#define PYL_DUP2D 2      /* 2-D Cython path: each weight deposited STENCIL times */
#define PYL_RENORM2D 4   /* ... then the whole grid divided by STENCIL */
#define PYL_ZERO 8       /* zero the grid first (MA itself accumulates) */
#define PYL_ATOMIC 16    /* force the direct global-atomic deposit */
#define PYL_BINNED 32    /* force the cell-binned LDS deposit */
#define PYL_SORTED 64    /* force the cell-binned LDS deposit after a presort by supertile (for
                            particles in no spatial order) */
#define PYL_ACC_F32W 1024  /* binned deposit into a float32 grid: 1-D weights and their products in
                              float32 from the float64 cell fraction, float64 cell sums (cells within
                              a few float32 ulps of the exact value; the same speed as the default) */
#define PYL_ACC_F64 2048   /* binned deposit (the default): float64 weights and float64 cell sums, so
                              every cell of a float32 grid is the correctly rounded exact value */

/* Pk flags */
#define PYL_PK_1D 1
#define PYL_PK_2D 2
#define PYL_PK_PHASE 4
#define PYL_PK_ALL 7
#define PYL_PK_ASYNC 256     /* no host synchronisation for the mode check (see pyl_pk) */
#define PYL_PK_DROP_KZ0 512  /* fault injection for tests of the mode check: the binning kernel leaves
                                out the kz = 0 plane (a normal run that loses modes; never set by the
                                library itself) */

int pyl_abi_version(void);
const char* pyl_last_error(void);

/* ---------------------------------------------------------------- device memory
 * No counterpart in the reference (its arrays live in host memory): the few device-memory calls a
 * caller without a device-array library needs to use the entry points below from host arrays (the
 * numpy drop-in path).  pyl_device_count reports 0 without a GPU or driver.  pyl_memcpy_h2d returns
 * once `src` may be reused: the data goes through the library's pinned staging buffers (3 x 32 MiB,
 * filled by host threads while the DMA of the previous chunk runs) and its copies are queued on
 * `stream` (NULL: the default stream).  pyl_memcpy_d2d and pyl_dev_zero are asynchronous on `stream`;
 * pyl_memcpy_d2h returns when the copy is done. */
int pyl_device_count(int* count);
int pyl_dev_alloc(void** ptr, int64_t bytes);
int pyl_dev_free(void* ptr);
int pyl_memcpy_h2d(void* dst, const void* src, int64_t bytes, void* stream);
int pyl_memcpy_d2h(void* dst, const void* src, int64_t bytes, void* stream);
int pyl_memcpy_d2d(void* dst, const void* src, int64_t bytes, void* stream);
int pyl_dev_zero(void* ptr, int64_t bytes, void* stream);
int pyl_stream_sync(void* stream);

/* ---------------------------------------------------------------- mass assignment
 * Replaces MAS_library.MA and the per-scheme Cython functions NGP/CIC/TSC/PCS, NGPW/CICW/TSCW/PCSW,
 * CICW_d/NGPW_d (MAS_library.pyx:57-545).  pos: (n, axes) float32 row-major; W: (n,) float32 or
 * NULL; grid: dims^axes float32 (float64 with PYL_GRID_F64), accumulated into (MA never zeroes,
 * MAS_library.pyx:159-166).  For the 2-D Cython semantics pass PYL_DUP2D [| PYL_RENORM2D]. */
int pyl_mas_deposit(const float* pos, const float* W, int64_t n, int axes, int dims, double BoxSize,
                    int scheme, int flags, void* grid, void* stream);

/* Overdensity of particles on a 3-D grid, delta = rho / rhobar - 1 with rhobar = M / dims^3 (M = n, or
 * the float64 sum of W): the deposit of pyl_pk_particles, written to `grid` (float32, float64 with
 * PYL_GRID_F64; overwritten), delta evaluated in float64 before one rounding.  Replaces MA followed by
 * `delta /= mean(delta); delta -= 1` (Documentation.md:117-118, Pk_snapshot.py:84-88).  Deposit flags
 * as for pyl_mas_deposit (PYL_ATOMIC / PYL_BINNED / PYL_SORTED, PYL_ACC_*); without them the deposit
 * is the autotuner's choice. */
int pyl_mas_overdensity(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme,
                        int flags, void* grid, void* stream);

/* Host-pointer drop-ins of the OpenMP C kernels, same argument list as MAS_c.h:3-10
 * (void NGP|CIC|TSC|PCS(FLOAT *pos, FLOAT *number, FLOAT *W, long particles, int dims, int axes,
 * FLOAT BoxSize, int threads)).  Synchronous; `threads` is accepted and ignored. */
int pyl_mas_NGP(float* pos, float* number, float* W, long particles, int dims, int axes, float BoxSize,
                int threads);
int pyl_mas_CIC(float* pos, float* number, float* W, long particles, int dims, int axes, float BoxSize,
                int threads);
int pyl_mas_TSC(float* pos, float* number, float* W, long particles, int dims, int axes, float BoxSize,
                int threads);
int pyl_mas_PCS(float* pos, float* number, float* W, long particles, int dims, int axes, float BoxSize,
                int threads);

/* ---------------------------------------------------------------- FFT
 * Replace FFT3Dr_f/_d (unnormalised r2c, (N,N,N) -> (N,N,N/2+1)) and IFFT3Dr_f/_d (c2r,
 * 1/N^3-normalised when normalise != 0), Pk_library.pyx:120-181.  The input of pyl_fft_c2r is
 * not modified. */
int pyl_fft_r2c(const void* in, void* out, int dims, int prec, void* stream);
int pyl_fft_c2r(const void* in, void* out, int dims, int prec, int normalise, void* stream);

/* ---------------------------------------------------------------- power spectrum
 * Output layout of pyl_pk / pyl_pk_particles (float64, out_len values):
 *   k3D[kmax] P0[kmax] P2[kmax] P4[kmax] Pkphase[kmax] Nmodes3D[kmax]
 *   k1D[kmax_par] Pk1D[kmax_par] Nmodes1D[kmax_par]
 *   kpar[n2] kper[n2] Pk2D[n2] Nmodes2D[n2],   n2 = (kmax_par+1)*(kmax_per+1)
 * i.e. the attributes of Pk_library.Pk (Pk_library.pyx:387-425). */
int pyl_pk_layout(int dims, int* kmax_par, int* kmax_per, int* kmax, int64_t* out_len);

/* Replaces Pk_library.Pk (Pk_library.pyx:266-425).  delta: (N,N,N) real of precision prec.
 * mas_p: window exponent (0 = no deconvolution, as MAS='None').  deltak_out (nullable) receives
 * the corrected half spectrum (keep_deltak, Pk_library.pyx:423).
 * Nmodes3D/1D/2D are the counts of the binning kernel itself (Pk_library.pyx:376-381).  Every bin's
 * count is compared on the device with the independent-mode geometry; the call then synchronises
 * `stream` and returns PYL_E_MODES when any bin differs -- which includes a total different from
 * (N^3-8)/2+8 (check_number_modes, Pk_library.pyx:90-102, 411).  With PYL_PK_ASYNC (or inside a
 * stream capture) the call does not synchronise and the caller checks the returned Nmodes. */
int pyl_pk(const void* delta, int dims, double BoxSize, int axis, int mas_p, int prec, int flags, double* out,
           void* deltak_out, void* stream);

/* No counterpart in the reference.  pyl_pk for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same arguments
 * and results, but every scratch buffer and lookup table of the call lives in a fixed per-GPU array
 * of PYL_NOALLOC_SCRATCH_MB MiB inside the library, so the call allocates nothing.  dims must be at
 * most PYL_NOALLOC_MAX_DIMS (PYL_E_UNSUPPORTED otherwise).  Every call on a GPU shares that GPU's
 * array and leaves its contents there: calls that overlap on one GPU (other host threads, or an
 * earlier call still running on another stream) race; host threads are serialised per device. */
#define PYL_NOALLOC_MAX_DIMS 128
#define PYL_NOALLOC_SCRATCH_MB 64
int pyl_pk_noalloc(const void* delta, int dims, double BoxSize, int axis, int mas_p, int prec, int flags, double* out,
                   void* deltak_out, void* stream);

/* Fused end-to-end path of Pk_snapshot.Pk_comp (Pk_snapshot.py:80-91): deposit with `scheme`,
 * delta = rho/rhobar - 1 with rhobar = sum(W)/N^3 (N_p/N^3 without weights), r2c, window
 * deconvolution with the same scheme, binning.  The density grid is not returned.  The mode check
 * and PYL_PK_ASYNC behave as in pyl_pk. */
int pyl_pk_particles(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, int axis,
                     int prec, int flags, double* out, void* stream);

/* Replaces Pk_library.correct_MAS (Pk_library.pyx:1749-1806): r2c of the (dims^3) field, every
 * independent mode multiplied by the float window factor prod_i (x_i/sin x_i)^mas_p (the dependent
 * half of the kz = 0 and N/2 planes unchanged, as the reference), 1/N^3-normalised c2r into `out`.
 * delta, out: device arrays of precision prec (may not alias). */
int pyl_correct_mas(const void* delta, void* out, int dims, int mas_p, int prec, void* stream);

/* Replaces Pk_library.Xi (Pk_library.pyx:2035-2151).  out: HOST float64 array of 5*kmax values
 * (pyl_xi_layout): r3D (kmax), xi (kmax x 3: monopole, quadrupole, hexadecapole, row-major),
 * Nmodes3D (kmax); bins floor(|r|) in cell units, r = 0 bin dropped.  Synchronises `stream`. */
int pyl_xi_layout(int dims, int* kmax, int64_t* out_len);
int pyl_xi(const void* delta, int dims, double BoxSize, int mas_p, int axis, int prec, double* out, void* stream);

/* Replace smoothing_library.FT_filter / field_smoothing (smoothing_library.pyx:20-114).
 * pyl_ft_filter: top-hat (gaussian = 0) or Gaussian filter of radius R on the dims^3 grid,
 * normalised by its float64 sum, r2c into out_k (complex64 (dims, dims, dims/2+1), device).
 * pyl_field_smoothing: r2c(field) * filter_k, 1/N^3-normalised c2r into out (float32, device). */
int pyl_ft_filter(int dims, float BoxSize, float R, int gaussian, void* out_k, void* stream);
int pyl_field_smoothing(const void* field, const void* filter_k, void* out, int dims, void* stream);

/* Overdensity and field combinations of the snapshot drivers, with numpy's float32 rounding:
 * pyl_field_div_sub: field = field / div - sub in place, div and sub rounded to the grid type and
 * each operation rounded once -- `delta /= mean; delta -= 1.0` (Pk_snapshot.py:88, 194) and
 * `delta_tot /= Omega_tot` (Pk_snapshot.py:256, sub = 0).
 * pyl_field_axpy: y = y + a * x with the product rounded before the sum -- `delta_tot +=
 * Omega * delta` (Pk_snapshot.py:252).  n cells of precision prec (PYL_F32 | PYL_F64), device. */
int pyl_field_div_sub(void* field, int64_t n, double div, double sub, int prec, void* stream);
int pyl_field_axpy(void* y, const void* x, int64_t n, double a, int prec, void* stream);

/* Redshift space fused into the same path: pos_redshift_space (redshift_space_library.pyx:29-43)
 * applied to each position as the deposit loads it -- x[rsd_axis] += vel[rsd_axis]*(1+z)/H in
 * float32, then (x + L) mod L when it leaves [0, L] -- followed by pyl_pk_particles.  `pos` is not
 * modified.  vel: (n, 3) float32 device array; Hubble in (km/s)/(Mpc/h).  The result equals
 * pyl_pk_particles on the positions pos_redshift_space would produce. */
int pyl_pk_particles_rsd(const float* pos, const float* vel, const float* W, int64_t n, int dims, double BoxSize,
                         float Hubble, float redshift, int rsd_axis, int scheme, int axis, int prec, int flags,
                         double* out, void* stream);

/* Replaces redshift_space_library.pos_redshift_space(pos, vel, BoxSize, Hubble, redshift, axis)
 * (redshift_space_library.pyx:29-43): in place, pos[i, axis] += vel[i, axis] * (float)((1+z)/H) in
 * float32 (product and sum rounded separately), then (pos + L) mod L (fmod) where the result
 * leaves [0, L].  pos, vel: (n, cols) float32 row-major device arrays, 0 <= axis < cols <= 3. */
int pyl_pos_redshift_space(float* pos, const float* vel, int64_t n, int cols, float BoxSize, float Hubble,
                           float redshift, int axis, void* stream);

/* Replaces Pk_library.XPk (Pk_library.pyx:534-798): auto and cross power spectra of nfields
 * (1..8) fields.  deltas: HOST array of nfields device pointers to (N,N,N) reals of precision prec;
 * mas_p: HOST array of window exponents.  out: HOST float64 array of pyl_xpk_layout values:
 *   k3D[kmax] Nmodes3D[kmax] k1D[kmax_par] Nmodes1D[kmax_par] kpar[n2] kper[n2] Nmodes2D[n2]
 *   then, over the C = F(F+1)/2 pairs (autos 0..F-1, then crosses (0,1),(0,2),...,(1,2),...) as the
 *   last, fastest index: P[kmax][3][C] (P0,P2,P4)  P1D[kmax_par][C]  P2D[n2][C].
 * The call synchronises `stream`. */
int pyl_xpk_layout(int dims, int nfields, int64_t* out_len);
int pyl_xpk(const void* const* deltas, int nfields, int dims, double BoxSize, int axis, const int* mas_p, int prec,
            double* out, void* stream);

/* ---------------------------------------------------------------- bispectrum
 * Replaces Pk_library.Bk (bispectrum_library.pyx:32-196).  theta: HOST array of ntheta angles.
 * B, Q (ntheta) and k, Pk (ntheta+2) are HOST float64 arrays; the call synchronises `stream`. */
int pyl_bk(const void* delta, int dims, double BoxSize, double k1, double k2, const double* theta, int ntheta,
           int mas_p, int prec, double* B, double* Q, double* k, double* Pk, void* stream);

/* ---------------------------------------------------------------- batches
 * Builder-defined (the reference processes one field per call, Pk/Pk.py:17-27).  A batch is an
 * array of independent particles -> P(k) jobs of any sizes, run back to back on one stream; the
 * jobs share the library workspaces.  pyl_graph_create runs the batch once inside a private set of
 * workspaces owned by the graph (sizing every buffer and plan), then captures it into a hipGraph;
 * pyl_graph_launch replays the whole batch with one launch.  Because the graph owns its workspaces,
 * eager calls made between replays (of any size) never invalidate it; launches of one graph must
 * not overlap each other (replay it on one stream).  pyl_graph_destroy synchronises the device and
 * frees the graph and its workspaces.  Inputs are read and outputs written at the device addresses
 * given in the jobs.  Jobs run without the host-side mode check (PYL_PK_ASYNC is implied): the
 * caller checks the Nmodes of each output. */
typedef struct {
    const float* pos;   /* (n, 3) float32, device */
    const float* W;     /* (n,) float32 or NULL */
    int64_t n;
    int dims;
    double BoxSize;
    int scheme;         /* 1..4, also the window exponent */
    int axis;
    int prec;           /* PYL_F32 / PYL_F64 */
    int flags;          /* PYL_PK_* | PYL_ATOMIC / PYL_BINNED / PYL_SORTED */
    double* out;        /* device, pyl_pk_layout(dims) values */
} pyl_pk_job;

int pyl_pk_batch(const pyl_pk_job* jobs, int njobs, void* stream);
int pyl_graph_create(const pyl_pk_job* jobs, int njobs, void* stream, void** graph);
int pyl_graph_launch(void* graph, void* stream);
int pyl_graph_destroy(void* graph);

/* ---------------------------------------------------------------- differentiation
 * Builder-defined: the reference has no adjoint.  Gradient of L = sum_{b<nb} w[b] P0[b] (P0 of
 * pyl_pk_particles with `scheme`, no weights, line of sight irrelevant for P0; bin b is the b-th
 * output bin, |k| in [b+1, b+2) kF) with respect to the particle positions.  w: device float64
 * (nb); grad: device float32 (n, 3); out: device float64 pyl_pk_layout vector of the forward pass
 * (P0 etc., so L can be formed).  Nearest reference forms: CIC_interp (MAS_library.pyx:558-599),
 * the transpose of the CIC scatter, and IFFT3Dr_f (Pk_library.pyx:152-165). */
int pyl_pk_adjoint(const float* pos, int64_t n, int dims, double BoxSize, int scheme, const double* w, int nb,
                   float* grad, double* out, void* stream);

/* Replaces MAS_library.CIC_interp(density, BoxSize, pos, den) (MAS_library.pyx:558-599) for scheme 2,
 * and generalises it to NGP/TSC/PCS: out[p] = sum over the stencil cells of density * weight.
 * density: device float32 (dims^3); pos: device float32 (n, 3); out: device float32 (n). */
int pyl_interp(const float* density, int dims, double BoxSize, const float* pos, int64_t n, int scheme, float* out,
               void* stream);

/* ---------------------------------------------------------------- deposit autotuner
 * Builder-defined.  Without PYL_ATOMIC / PYL_BINNED / PYL_SORTED, pyl_pk_particles picks the
 * deposit kernel per (dims, scheme, prec, weights, floor(log2 n)) by timing the three once (outside
 * graph capture) and caching the fastest.  pyl_tune(0) disables timing (binned is used where
 * supported); pyl_tune_export / pyl_tune_import move the cache as text lines
 * "dims scheme prec hasw nlog choice" (choice 0 = binned, 1 = atomic, 2 = presorted binned);
 * export returns the full text length. */
int pyl_tune(int enable);
int pyl_tune_export(char* buf, int len);
int pyl_tune_import(const char* text);

/* ---------------------------------------------------------------- instrumentation
 * Builder-defined (no counterpart in the reference, whose only timing is printed wall time,
 * MAS_library.pyx:71,112; Pk_library.pyx:269,382,425).  pyl_profile(1) resets the accumulators and
 * records a hipEvent pair around every stage on the stream the stage runs on; pyl_profile_read
 * writes "stage count total_ms" lines into buf (synchronising on the recorded events) and returns
 * the full length of the text. */
int pyl_profile(int enable);
int pyl_profile_read(char* buf, int len);

#ifdef __cplusplus
}
#endif

#endif /* PYLIANS_HIP_H */
