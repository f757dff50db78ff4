// Power-spectrum estimator pieces (Pk_library.pyx:266-425 semantics) for gfx950.
#pragma once
#include "common.hpp"

namespace pyl {

struct PkDims {
    int N, middle, kmax_par, kmax_per, kmax;
    int64_t P;                 // half-spectrum row pitch (complex elements)
    int n1() const { return kmax_par + 1; }
    int n2() const { return (kmax_par + 1) * (kmax_per + 1); }
    int n3() const { return kmax + 1; }
};
PkDims pk_dims(int N, int64_t P);

// Data-independent bin geometry of the independent modes (Pk_library.pyx:322-330):
// integer mode counts per 1-D / 2-D / 3-D bin and the float64 sum of |k| per 3-D bin
// (kF units).  Computed once per (device, N, axis) on the GPU and cached.
struct PkGeometry {
    PkDims d;
    int axis;
    // device arrays
    unsigned long long* n1;    // n1()
    unsigned long long* n2;    // n2()
    unsigned long long* n3;    // n3()
    double* k3sum;             // n3()
    // host copies
    std::vector<unsigned long long> h_n1, h_n2, h_n3;
    std::vector<double> h_k3sum;
};
const PkGeometry* pk_geometry(int N, int axis, int* rc);

// float64 per-bin sums of the data-dependent quantities and the mode counts of the binning
// kernel (integers held in float64, exact below 2^53), device arrays owned by caller
struct PkSums {
    double* s3;   // n3 * 4: sum d2, sum d2 L2(mu), sum d2 L4(mu), sum phase^2
    double* s1;   // n1: sum d2 over modes with |k| <= N/2, binned by |k_par|
    double* s2;   // n2: sum d2 over 2-D bins (kmax_par+1)*k_per + k_par
    double* c3;   // n3: modes binned per 3-D bin
    double* c1;   // n1: modes binned per 1-D bin
    double* c2;   // n2: modes binned per 2-D bin
    unsigned* status;   // [1]: bins whose count differs from the mode geometry (set by pk_finalize)
};
size_t pk_sums_doubles(const PkDims& d);   // doubles of one PkSums block (status word included)
void pk_sums_carve(const PkDims& d, double* base, PkSums* s);

// PK_ASYNC: no host synchronisation for the mode check (the counts are still in the output).
// PK_DROP_KZ0: the binning kernel leaves out the kz = 0 plane -- fault injection for the tests of
// the mode check, never set by the library itself.
enum PkFlags { PK_1D = 1, PK_2D = 2, PK_PHASE = 4, PK_ASYNC = 256, PK_DROP_KZ0 = 512 };

// Axis of the last forward FFT pass for a line of sight `axis`: 1 (y) for axis 0, else 0 (x).
// The half spectrum must have gone through pass Z and the column pass along the other axis.
int pk_last_axis(int axis);
// That last pass fused with the window deconvolution and the binning of the independent modes.
// keep (nullable) receives the window-corrected spectrum (Pk_library.pyx:354-355, 423).
template <typename R>
int pk_last_pass_bin(const cplx<R>* spec, cplx<R>* keep, int N, int64_t P, int axis, int mas_p, int flags,
                     const PkSums& sums, hipStream_t st);

// Normalisation of Pk_library.pyx:387-421 on the device, with Nmodes from the binning kernel's
// counts; every bin whose count differs from the geometry is counted into sums.status (the
// check_number_modes test, Pk_library.pyx:90-102, made per bin).  out layout (float64):
//   [k3D (kmax) | P0 (kmax) | P2 (kmax) | P4 (kmax) | Pkphase (kmax) | Nmodes3D (kmax) |
//    k1D (kmax_par) | Pk1D (kmax_par) | Nmodes1D (kmax_par) |
//    kpar (n2) | kper (n2) | Pk2D (n2) | Nmodes2D (n2)]
size_t pk_out_len(const PkDims& d);
int pk_finalize(const PkGeometry* g, const PkSums& sums, double BoxSize, int flags, double* out, hipStream_t st);

// Argument blocks of the fused last pass + binning, of the two reductions of its per-block
// histograms and of the normalisation (shared with the batch planner)
template <typename R> struct LastBinArgs {
    const cplx<R>* src;
    cplx<R>* keep;
    long long P;
    int W, last, axis, fixk, nchunk, flags, kmax_par, kmax_per, n1, n3;
    const double* win;
    const cplx<R>* tw;
    double* part;
    int RS;
};

struct ReduceArgs {
    const double* part;
    int RS, nblocks, n3, n1, n2w, nper, kmax_par, fixk, nchunk, NL, N;
    double *g3, *g1, *gc3, *gc1, *g2, *gc2;
};

struct FinArgs {
    int N, kmax, kmax_par, n2;
    double BoxSize;
    int flags;
    const unsigned long long *n1c, *n2c, *n3c;
    const double* k3s;
    const double *s3, *s1, *s2, *c3, *c1, *c2;
    unsigned* status;
    double* out;
};

// Batched last pass + binning + reductions + normalisation over jobs of one grid size N and line
// of sight class (z or not): *_job builds a job's arguments (part: its per-block histogram rows,
// last_bin_part_doubles(N, axis, prec) of them) and block count; last_bin_batch launches the
// tables (block ranges start[0..nj]).
size_t last_bin_part_doubles(int N, int axis, int prec);
template <typename R>
int last_bin_job(const cplx<R>* spec, int N, int64_t P, int axis, int mas_p, int flags, const PkSums& sums,
                 double* part, LastBinArgs<R>* a, ReduceArgs* r, unsigned* blocks);
template <typename R>
int last_bin_batch(int N, int axis, const LastBinArgs<R>* tab, const unsigned* start, int nj, unsigned total,
                   const ReduceArgs* rtab, hipStream_t st);
FinArgs finalize_args(const PkGeometry* g, const PkSums& sums, double BoxSize, int flags, double* out);
int finalize_batch(int N, const FinArgs* tab, int nj, hipStream_t st);

// Deposit algorithm of the fused particles -> P(k) path: cell-binned, direct atomic, or cell-binned
// after a presort by supertile.  deposit_kind: the one `flags` forces (DEP_ATOMIC / DEP_BINNED /
// DEP_SORTED), else the autotuner's (timing trials the first time a size class is seen, which
// overwrite `grid`).  pk_particles: the whole fused path of one job (pyl_pk_particles).
enum DepositKind { DK_BINNED = 0, DK_ATOMIC = 1, DK_SORTED = 2 };
struct Rsd;
int deposit_kind(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, bool f64,
                 int flags, void* grid, const double* dmass, hipStream_t st, int* kind, const Rsd& rsd);
int pk_particles(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, int axis,
                 int prec, int flags, double* out, hipStream_t st, const Rsd& rsd);

// The whole estimator from a real (N,N,N) field: pass Z r2c (affine: delta = rho * N^3 / mass - 1
// applied on the load, mass = dmass[0] when dmass is non-null), the column pass, the fused last pass
// + binning, the normalisation and the mode check (see pyl_pk).  keep (nullable) receives the
// corrected spectrum.
template <typename R>
int pk_from_delta(const R* delta, int dims, double BoxSize, int axis, int mas_p, int flags, double* out,
                  cplx<R>* keep, bool affine, const double* dmass, double mass, hipStream_t st);

}  // namespace pyl
