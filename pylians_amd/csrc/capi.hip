// C ABI of libpylians_hip (declared in include/pylians_hip.h).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "../../include/pylians_hip.h"
#include "adjoint.hpp"
#include "common.hpp"
#include "deposit.hpp"
#include "fft.hpp"
#include "pk.hpp"
#include "profile.hpp"
#include "util.hpp"

using namespace pyl;

static_assert(PYL_GRID_F64 == DEP_GRID_F64 && PYL_DUP2D == DEP_DUP2D && PYL_RENORM2D == DEP_RENORM2D &&
                  PYL_ZERO == DEP_ZERO && PYL_ATOMIC == DEP_ATOMIC && PYL_BINNED == DEP_BINNED &&
                  PYL_SORTED == DEP_SORTED && PYL_ACC_F32W == DEP_ACC_F32W && PYL_ACC_F64 == DEP_ACC_F64,
              "flag mismatch");
static_assert(PYL_PK_1D == PK_1D && PYL_PK_2D == PK_2D && PYL_PK_PHASE == PK_PHASE && PYL_PK_ASYNC == PK_ASYNC &&
                  PYL_PK_DROP_KZ0 == PK_DROP_KZ0,
              "flag mismatch");

namespace pyl {
int weight_sum(const float* W, int64_t n, double* out, hipStream_t st);
const double* pk_window_table(int N, int p);
}

extern "C" int pyl_abi_version(void) { return PYL_ABI_VERSION; }
extern "C" const char* pyl_last_error(void) { return last_error(); }

// ------------------------------------------------------------------ mass assignment
extern "C" int pyl_mas_deposit(const float* pos, const float* W, int64_t n, int axes, int dims, double BoxSize,
                               int scheme, int flags, void* grid, void* stream) {
    PYL_CHECK(n >= 0, E_ARG, "negative particle count");
    PYL_CHECK(n == 0 || pos, E_ARG, "null positions");
    PYL_CHECK(grid, E_ARG, "null grid");
    return deposit(pos, W, n, axes, dims, BoxSize, scheme, flags, grid, (hipStream_t)stream);
}

static int mas_c_host(int scheme, const float* pos, float* number, const float* W, long particles, int dims,
                      int axes, float BoxSize) {
    PYL_CHECK(particles >= 0 && dims > 0 && (axes == 2 || axes == 3), E_ARG, "bad arguments");
    PYL_CHECK(pos && number, E_ARG, "null pointer");
    size_t cells = (size_t)dims * dims * (axes == 3 ? dims : 1);
    int rc = OK;
    float* dpos = (float*)workspace(WS_HOST_POS, sizeof(float) * (size_t)particles * axes + 16, &rc);
    if (rc) return rc;
    float* dgrid = (float*)workspace(WS_HOST_GRID, sizeof(float) * cells, &rc);
    if (rc) return rc;
    float* dW = nullptr;
    if (W) {
        dW = (float*)workspace(WS_HOST_W, sizeof(float) * (size_t)particles + 16, &rc);
        if (rc) return rc;
    }
    PYL_HIP(hipMemcpy(dpos, pos, sizeof(float) * (size_t)particles * axes, hipMemcpyHostToDevice));
    PYL_HIP(hipMemcpy(dgrid, number, sizeof(float) * cells, hipMemcpyHostToDevice));
    if (W) PYL_HIP(hipMemcpy(dW, W, sizeof(float) * (size_t)particles, hipMemcpyHostToDevice));
    // MAS_c.c 2-D path deposits every weight once (n_max = 1): no DUP2D
    rc = deposit(dpos, dW, particles, axes, dims, (double)BoxSize, scheme, 0, dgrid, 0);
    if (rc) return rc;
    PYL_HIP(hipMemcpy(number, dgrid, sizeof(float) * cells, hipMemcpyDeviceToHost));
    return OK;
}

extern "C" int pyl_mas_NGP(float* pos, float* number, float* W, long particles, int dims, int axes, float BoxSize,
                           int threads) {
    (void)threads;
    return mas_c_host(1, pos, number, W, particles, dims, axes, BoxSize);
}
extern "C" int pyl_mas_CIC(float* pos, float* number, float* W, long particles, int dims, int axes, float BoxSize,
                           int threads) {
    (void)threads;
    return mas_c_host(2, pos, number, W, particles, dims, axes, BoxSize);
}
extern "C" int pyl_mas_TSC(float* pos, float* number, float* W, long particles, int dims, int axes, float BoxSize,
                           int threads) {
    (void)threads;
    return mas_c_host(3, pos, number, W, particles, dims, axes, BoxSize);
}
extern "C" int pyl_mas_PCS(float* pos, float* number, float* W, long particles, int dims, int axes, float BoxSize,
                           int threads) {
    (void)threads;
    return mas_c_host(4, pos, number, W, particles, dims, axes, BoxSize);
}

// ------------------------------------------------------------------ FFT
static int check_fft(int dims, int prec) {
    PYL_CHECK(prec == PYL_F32 || prec == PYL_F64, E_ARG, "prec must be 4 or 8");
    PYL_CHECK(fft_size_supported(dims), E_UNSUPPORTED,
              "grid size %d not supported (supported: 2^a 3^b 5^c, 8 <= N <= 1024, from the built-in list)", dims);
    return OK;
}

extern "C" int pyl_fft_r2c(const void* in, void* out, int dims, int prec, void* stream) {
    int rc = check_fft(dims, prec);
    if (rc) return rc;
    const int64_t P = half_len(dims);
    hipStream_t st = (hipStream_t)stream;
    if (prec == PYL_F32) return fft3d_r2c<float>((const float*)in, (cplx<float>*)out, dims, P, st);
    return fft3d_r2c<double>((const double*)in, (cplx<double>*)out, dims, P, st);
}

extern "C" int pyl_fft_c2r(const void* in, void* out, int dims, int prec, int normalise, void* stream) {
    int rc = check_fft(dims, prec);
    if (rc) return rc;
    const int64_t P = half_len(dims);
    const size_t nspec = (size_t)dims * dims * P;
    hipStream_t st = (hipStream_t)stream;
    const double sc = normalise ? 1.0 / ((double)dims * dims * dims) : 1.0;
    void* ws = workspace(WS_SPEC2, nspec * (prec == PYL_F32 ? 8 : 16), &rc);
    if (rc) return rc;
    if (prec == PYL_F32)
        return fft3d_c2r<float>((const cplx<float>*)in, (float*)out, (cplx<float>*)ws, dims, P, (float)sc, st);
    return fft3d_c2r<double>((const cplx<double>*)in, (double*)out, (cplx<double>*)ws, dims, P, sc, st);
}

// ------------------------------------------------------------------ power spectrum
extern "C" int pyl_pk_layout(int dims, int* kmax_par, int* kmax_per, int* kmax, int64_t* out_len) {
    PYL_CHECK(dims > 1, E_ARG, "dims must be > 1");
    PkDims d = pk_dims(dims, half_len(dims));
    if (kmax_par) *kmax_par = d.kmax_par;
    if (kmax_per) *kmax_per = d.kmax_per;
    if (kmax) *kmax = d.kmax;
    if (out_len) *out_len = (int64_t)pk_out_len(d);
    return OK;
}

// float64 sums and mode counts live in one zeroed workspace (pk_sums_carve)
static int pk_sums_ws(const PkDims& d, PkSums* s, hipStream_t st) {
    int rc = OK;
    const size_t n = pk_sums_doubles(d);
    double* b = (double*)workspace(WS_SUMS, sizeof(double) * n, &rc);
    if (rc) return rc;
    rc = zero_fill(b, sizeof(double) * n, st);
    if (rc) return rc;
    pk_sums_carve(d, b, s);
    return OK;
}

// check_number_modes (Pk_library.pyx:90-102) on the counts of the binning kernel: pk_finalize
// counted the bins whose count differs from the independent-mode geometry into sums.status.
// Synchronises the stream unless PK_ASYNC is set or the stream is being captured (then the
// counts are only in the output's Nmodes, for the caller to check).
static int check_modes(const PkGeometry* g, const PkSums& sums, int flags, hipStream_t st) {
    if (flags & PK_ASYNC) return OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    PYL_HIP(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) return OK;
    unsigned bad = 0u;
    PYL_HIP(hipMemcpyAsync(&bad, sums.status, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    PYL_HIP(hipStreamSynchronize(st));
    if (bad == 0u) return OK;
    const PkDims& d = g->d;
    std::vector<double> c3(d.n3());
    PYL_HIP(hipMemcpy(c3.data(), sums.c3, sizeof(double) * d.n3(), hipMemcpyDeviceToHost));
    double tot = 0.0;
    for (double v : c3) tot += v;
    const long long N = d.N;
    const long long own = (N % 2 == 1) ? 1 : 8;
    const long long expect = (N * N * N - own) / 2 + own;
    set_error("Not all modes counted: counted %.0f independent modes, expected %lld (%u bins differ from the "
              "mode geometry)", tot, expect, bad);
    return E_MODES;
}

namespace pyl {
template <typename R>
int pk_from_delta(const R* delta, int dims, double BoxSize, int axis, int mas_p, int flags, double* out,
                  cplx<R>* keep, bool affine, const double* dmass, double mass, hipStream_t st) {
    int rc = OK;
    // rows of the internal spectrum padded to 128 bytes: the strided column passes then read whole
    // cache lines (kz = N/2+1 .. P-1 are never touched)
    constexpr int64_t ALIGN = 128 / (int64_t)sizeof(cplx<R>);
    const int64_t P = (half_len(dims) + ALIGN - 1) / ALIGN * ALIGN;
    const PkGeometry* g = pk_geometry(dims, axis, &rc);
    if (rc) return rc;
    cplx<R>* spec = (cplx<R>*)workspace(WS_SPEC, sizeof(cplx<R>) * (size_t)dims * dims * P, &rc);
    if (rc) return rc;
    PkSums sums;
    rc = pk_sums_ws(g->d, &sums, st);
    if (rc) return rc;
    {
        StageTimer t("fft_z_r2c", st);
        rc = fft_pass_z_r2c<R>(delta, spec, dims, P, affine, dmass, mass, st);
        if (rc) return rc;
    }
    const int last = pk_last_axis(axis);
    {
        StageTimer t("fft_mid", st);
        rc = fft_pass_cols<R>(spec, spec, dims, P, last == 0 ? 1 : 0, false, (R)1, st);
        if (rc) return rc;
    }
    {
        StageTimer t("fft_last_bin", st);
        rc = pk_last_pass_bin<R>(spec, keep, dims, P, axis, mas_p, flags, sums, st);
        if (rc) return rc;
    }
    {
        StageTimer t("finalize", st);
        rc = pk_finalize(g, sums, BoxSize, flags, out, st);
        if (rc) return rc;
    }
    return check_modes(g, sums, flags, st);
}
template int pk_from_delta<float>(const float*, int, double, int, int, int, double*, cplx<float>*, bool, const double*,
                                  double, hipStream_t);
template int pk_from_delta<double>(const double*, int, double, int, int, int, double*, cplx<double>*, bool,
                                   const double*, double, hipStream_t);
}  // namespace pyl

extern "C" int pyl_pk(const void* delta, int dims, double BoxSize, int axis, int mas_p, int prec, int flags,
                      double* out, void* deltak_out, void* stream) {
    int rc = check_fft(dims, prec);
    if (rc) return rc;
    PYL_CHECK(axis >= 0 && axis <= 2, E_ARG, "axis must be 0, 1 or 2");
    PYL_CHECK(mas_p >= 0 && mas_p <= 4, E_ARG, "window exponent must be in [0, 4]");
    PYL_CHECK(BoxSize > 0, E_ARG, "BoxSize must be positive");
    PYL_CHECK(delta && out, E_ARG, "null pointer");
    hipStream_t st = (hipStream_t)stream;
    if (prec == PYL_F32)
        return pk_from_delta<float>((const float*)delta, dims, BoxSize, axis, mas_p, flags, out,
                                    (cplx<float>*)deltak_out, false, nullptr, 1.0, st);
    return pk_from_delta<double>((const double*)delta, dims, BoxSize, axis, mas_p, flags, out,
                                 (cplx<double>*)deltak_out, false, nullptr, 1.0, st);
}

// ------------------------------------------------------------------ deposit autotuner
// Per (device, dims, scheme, precision, weights, floor(log2 n)) the fused pipeline times the
// cell-binned deposit, the direct-atomic deposit and the binned deposit after a presort by
// supertile once, after one untimed run of each (all overwrite the grid, so the trial runs are
// harmless), keeps the fastest and caches the choice.  pyl_tune_export / pyl_tune_import move the cache as text ("dims scheme prec
// hasw nlog choice" lines; choice 0 = binned, 1 = atomic, 2 = presorted binned).
namespace {
struct TuneKey {
    int dev, dims, scheme, prec, hasw, nlog;
    bool operator<(const TuneKey& o) const {
        return std::tie(dev, dims, scheme, prec, hasw, nlog) < std::tie(o.dev, o.dims, o.scheme, o.prec, o.hasw, o.nlog);
    }
};
std::mutex g_tmu;
std::map<TuneKey, int> g_tune;
bool g_tune_on = true;
}  // namespace

// Writes the overdensity into `grid` except on the float64 atomic path, which leaves rho there
// (pass Z then applies delta = rho * a - 1 in float64).  The float32 atomic path accumulates in a
// float64 grid and rounds delta once, like the binned path, so both give the same float32 field.
static int deposit_delta(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, bool f64,
                         int kind, void* grid, const double* dmass, hipStream_t st, const Rsd& rsd = Rsd(),
                         int flags = 0) {
    if (kind != DK_ATOMIC)
        return deposit_binned(pos, W, n, dims, BoxSize, scheme, f64, 2, 1.0, grid, dmass, (double)n, st, rsd,
                              kind == DK_SORTED, acc_mode(flags));
    if (f64) return deposit(pos, W, n, 3, dims, BoxSize, scheme, DEP_ZERO | DEP_ATOMIC | DEP_GRID_F64, grid, st, rsd);
    const int64_t cells = (int64_t)dims * dims * dims;
    int rc = OK;
    double* rho = (double*)workspace(WS_AUX, sizeof(double) * (size_t)cells, &rc);
    if (rc) return rc;
    rc = deposit(pos, W, n, 3, dims, BoxSize, scheme, DEP_ZERO | DEP_ATOMIC | DEP_GRID_F64, rho, st, rsd);
    if (rc) return rc;
    return delta_round(rho, (float*)grid, cells, dmass, (double)n, st);
}

static int tune_choice(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, bool f64,
                       void* grid, const double* dmass, hipStream_t st, int* kind, const Rsd& rsd = Rsd(),
                       int flags = 0) {
    if (!binned_supported(3, dims)) {
        *kind = DK_ATOMIC;
        return OK;
    }
    int dev = 0;
    PYL_HIP(hipGetDevice(&dev));
    int nlog = 0;
    while ((1LL << (nlog + 1)) <= n) ++nlog;
    TuneKey key{dev, dims, scheme, f64 ? 8 : (4 + acc_mode(flags) * 16), W ? 1 : 0, nlog};
    {
        std::lock_guard<std::mutex> lk(g_tmu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            *kind = it->second;
            return OK;
        }
        if (!g_tune_on) {
            *kind = DK_BINNED;
            return OK;
        }
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    PYL_HIP(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) {   // never time inside a capture
        *kind = DK_BINNED;
        return OK;
    }
    hipEvent_t e[2];
    for (auto& x : e) PYL_HIP(hipEventCreate(&x));
    float ms[3] = {0.f, 0.f, 0.f};
    int rc = OK;
    // (a prefix of the particles would not do: in spatially ordered inputs it is a slab of the box,
    // whose tiles hold several times the particles the bucket capacities are sized for)
    const int64_t ns = n;
    for (int c = 0; c < 3 && rc == OK; ++c) {
        rc = deposit_delta(pos, W, ns, dims, BoxSize, scheme, f64, c, grid, dmass, st, rsd, flags);  // warm
        if (rc) break;
        PYL_HIP(hipEventRecord(e[0], st));
        rc = deposit_delta(pos, W, ns, dims, BoxSize, scheme, f64, c, grid, dmass, st, rsd, flags);
        PYL_HIP(hipEventRecord(e[1], st));
        PYL_HIP(hipEventSynchronize(e[1]));
        PYL_HIP(hipEventElapsedTime(&ms[c], e[0], e[1]));
    }
    for (auto& x : e) (void)hipEventDestroy(x);
    if (rc) return rc;
    int best = DK_BINNED;
    for (int c = 1; c < 3; ++c)
        if (ms[c] < ms[best]) best = c;
    *kind = best;
    std::lock_guard<std::mutex> lk(g_tmu);
    g_tune[key] = best;
    return OK;
}

extern "C" int pyl_tune(int enable) {
    std::lock_guard<std::mutex> lk(g_tmu);
    g_tune_on = enable != 0;
    return OK;
}

extern "C" int pyl_tune_export(char* buf, int len) {
    std::lock_guard<std::mutex> lk(g_tmu);
    std::string s;
    for (auto& kv : g_tune) {
        char line[128];
        snprintf(line, sizeof(line), "%d %d %d %d %d %d\n", kv.first.dims, kv.first.scheme, kv.first.prec,
                 kv.first.hasw, kv.first.nlog, kv.second);
        s += line;
    }
    if (buf && len > 0) snprintf(buf, (size_t)len, "%s", s.c_str());
    return (int)s.size();
}

extern "C" int pyl_tune_import(const char* text) {
    PYL_CHECK(text, E_ARG, "null text");
    int dev = 0;
    PYL_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_tmu);
    const char* p = text;
    int d, sc, pr, hw, nl, ch, used;
    while (sscanf(p, "%d %d %d %d %d %d%n", &d, &sc, &pr, &hw, &nl, &ch, &used) == 6) {
        g_tune[TuneKey{dev, d, sc, pr, hw, nl}] = ch;
        p += used;
    }
    return OK;
}

namespace pyl {
int deposit_kind(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, bool f64,
                 int flags, void* grid, const double* dmass, hipStream_t st, int* kind, const Rsd& rsd) {
    if (flags & DEP_ATOMIC) {
        *kind = DK_ATOMIC;
        return OK;
    }
    if (flags & (DEP_BINNED | DEP_SORTED)) {
        PYL_CHECK(binned_supported(3, dims), E_UNSUPPORTED, "binned deposit needs dims %% 16 == 0");
        *kind = (flags & DEP_SORTED) ? DK_SORTED : DK_BINNED;
        return OK;
    }
    return tune_choice(pos, W, n, dims, BoxSize, scheme, f64, grid, dmass, st, kind, rsd, flags);
}

int pk_particles(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, int axis,
                 int prec, int flags, double* out, hipStream_t st, const Rsd& rsd) {
    int rc = check_fft(dims, prec);
    if (rc) return rc;
    PYL_CHECK(axis >= 0 && axis <= 2, E_ARG, "axis must be 0, 1 or 2");
    PYL_CHECK(scheme >= 1 && scheme <= 4, E_ARG, "unknown mass-assignment scheme %d", scheme);
    PYL_CHECK(n > 0 && pos && out, E_ARG, "bad particle arguments");
    const size_t cells = (size_t)dims * dims * dims;
    const bool f64 = prec == PYL_F64;
    void* grid = workspace(WS_GRID, cells * (f64 ? 8 : 4), &rc);
    if (rc) return rc;
    double* dmass = nullptr;
    if (W) {
        dmass = (double*)workspace(WS_MASS, 64, &rc);
        if (rc) return rc;
        rc = weight_sum(W, n, dmass, st);
        if (rc) return rc;
    }
    // binned path: the deposit writes delta = rho/rhobar - 1 itself (float64 before one rounding);
    // atomic path: rho, normalised inside pass Z
    int kind;
    rc = deposit_kind(pos, W, n, dims, BoxSize, scheme, f64, flags, grid, dmass, st, &kind, rsd);
    if (rc) return rc;
    rc = deposit_delta(pos, W, n, dims, BoxSize, scheme, f64, kind, grid, dmass, st, rsd, flags);
    if (rc) return rc;
    const int pkflags = flags & (PYL_PK_ALL | PYL_PK_ASYNC | PYL_PK_DROP_KZ0);
    if (f64)
        return pk_from_delta<double>((const double*)grid, dims, BoxSize, axis, scheme, pkflags, out, nullptr,
                                     kind == DK_ATOMIC, dmass, (double)n, st);
    return pk_from_delta<float>((const float*)grid, dims, BoxSize, axis, scheme, pkflags, out, nullptr, false,
                                dmass, (double)n, st);
}
}  // namespace pyl

// The deposit of pyl_pk_particles without the transform: delta = rho / rhobar - 1 on a 3-D grid.
extern "C" int pyl_mas_overdensity(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme,
                                   int flags, void* grid, void* stream) {
    PYL_CHECK(n > 0 && pos && grid, E_ARG, "bad particle or grid arguments");
    PYL_CHECK(dims > 0 && BoxSize > 0, E_ARG, "dims and BoxSize must be positive");
    PYL_CHECK(scheme >= 1 && scheme <= 4, E_ARG, "unknown mass-assignment scheme %d", scheme);
    hipStream_t st = (hipStream_t)stream;
    const bool f64 = flags & PYL_GRID_F64;
    int rc = OK;
    double* dmass = nullptr;
    if (W) {
        dmass = (double*)workspace(WS_MASS, 64, &rc);
        if (rc) return rc;
        rc = weight_sum(W, n, dmass, st);
        if (rc) return rc;
    }
    int kind;
    rc = deposit_kind(pos, W, n, dims, BoxSize, scheme, f64, flags, grid, dmass, st, &kind, Rsd());
    if (rc) return rc;
    rc = deposit_delta(pos, W, n, dims, BoxSize, scheme, f64, kind, grid, dmass, st, Rsd(), flags);
    if (rc) return rc;
    // the float64 atomic deposit leaves rho in the grid
    if (f64 && kind == DK_ATOMIC) return delta_inplace((double*)grid, (int64_t)dims * dims * dims, dmass, (double)n, st);
    return OK;
}

extern "C" int pyl_pk_particles(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme,
                                int axis, int prec, int flags, double* out, void* stream) {
    return pk_particles(pos, W, n, dims, BoxSize, scheme, axis, prec, flags, out, (hipStream_t)stream, Rsd());
}

extern "C" int pyl_pk_particles_rsd(const float* pos, const float* vel, const float* W, int64_t n, int dims,
                                    double BoxSize, float Hubble, float redshift, int rsd_axis, int scheme, int axis,
                                    int prec, int flags, double* out, void* stream) {
    PYL_CHECK(vel, E_ARG, "null velocity array");
    PYL_CHECK(rsd_axis >= 0 && rsd_axis <= 2, E_ARG, "rsd_axis must be 0, 1 or 2");
    PYL_CHECK(Hubble != 0.0f, E_ARG, "Hubble must be non-zero");
    Rsd r;
    r.vel = vel;
    r.axis = rsd_axis;
    r.fac = (float)((1.0 + (double)redshift) / (double)Hubble);   // redshift_space_library.pyx:36
    r.box = (float)BoxSize;
    return pk_particles(pos, W, n, dims, BoxSize, scheme, axis, prec, flags, out, (hipStream_t)stream, r);
}

extern "C" int pyl_pos_redshift_space(float* pos, const float* vel, int64_t n, int cols, float BoxSize, float Hubble,
                                      float redshift, int axis, void* stream) {
    PYL_CHECK(n >= 0, E_ARG, "negative particle count");
    PYL_CHECK(n == 0 || (pos && vel), E_ARG, "null position or velocity array");
    PYL_CHECK(cols >= 1 && cols <= 3 && axis >= 0 && axis < cols, E_ARG, "axis %d outside the %d columns", axis, cols);
    PYL_CHECK(Hubble != 0.0f, E_ARG, "Hubble must be non-zero");
    const float fac = (float)((1.0 + (double)redshift) / (double)Hubble);   // redshift_space_library.pyx:36
    return pos_redshift_space(pos, vel, n, cols, axis, fac, BoxSize, (hipStream_t)stream);
}

// ------------------------------------------------------------------ adjoint and interpolation
extern "C" int pyl_pk_adjoint(const float* pos, int64_t n, int dims, double BoxSize, int scheme, const double* w,
                              int nb, float* grad, double* out, void* stream) {
    int rc = check_fft(dims, PYL_F32);
    if (rc) return rc;
    PYL_CHECK(scheme >= 1 && scheme <= 4, E_ARG, "unknown mass-assignment scheme %d", scheme);
    PYL_CHECK(n > 0 && pos && w && grad && out, E_ARG, "bad arguments");
    const PkDims d = pk_dims(dims, half_len(dims));
    PYL_CHECK(nb >= 0 && nb <= d.kmax, E_ARG, "nb must be in [0, kmax = %d]", d.kmax);
    hipStream_t st = (hipStream_t)stream;
    const int64_t P = half_len(dims);
    const size_t cells = (size_t)dims * dims * dims;
    float* grid = (float*)workspace(WS_GRID, cells * 4, &rc);
    if (rc) return rc;
    cplx<float>* keep = (cplx<float>*)workspace(WS_SPEC2, sizeof(cplx<float>) * (size_t)dims * dims * P, &rc);
    if (rc) return rc;
    // forward: delta, spectrum (kept, window-corrected) and the binned outputs (out = P0 etc.)
    const int kind = binned_supported(3, dims) ? DK_BINNED : DK_ATOMIC;
    {
        StageTimer t("adj_forward_deposit", st);
        rc = deposit_delta(pos, nullptr, n, dims, BoxSize, scheme, false, kind, grid, nullptr, st);
        if (rc) return rc;
    }
    rc = pk_from_delta<float>(grid, dims, BoxSize, 0, scheme, PK_ASYNC, out, keep, false, nullptr, (double)n, st);
    if (rc) return rc;
    const PkGeometry* g = pk_geometry(dims, 0, &rc);
    if (rc) return rc;
    const double* win = pk_window_table(dims, scheme);
    PYL_CHECK(win, E_ALLOC, "window table allocation failed");
    const double rhobar = (double)n / ((double)dims * dims * dims);
    {
        StageTimer t("adj_grad_spectrum", st);
        rc = grad_spectrum<float>(keep, dims, P, w, nb, g->n3, BoxSize, win, 1.0 / rhobar, st);
        if (rc) return rc;
    }
    {
        StageTimer t("adj_c2r", st);
        cplx<float>* ws = (cplx<float>*)workspace(WS_SPEC, sizeof(cplx<float>) * (size_t)dims * dims * P, &rc);
        if (rc) return rc;
        rc = fft3d_c2r<float>(keep, grid, ws, dims, P, 1.0f, st);
        if (rc) return rc;
    }
    StageTimer t("adj_gather", st);
    return gather_grad(grid, dims, BoxSize, pos, n, scheme, 1.0, grad, st);
}

extern "C" int pyl_interp(const float* density, int dims, double BoxSize, const float* pos, int64_t n, int scheme,
                          float* out, void* stream) {
    PYL_CHECK(dims > 0 && BoxSize > 0 && density && (n == 0 || (pos && out)), E_ARG, "bad arguments");
    PYL_CHECK(scheme >= 1 && scheme <= 4, E_ARG, "unknown mass-assignment scheme %d", scheme);
    return interp(density, dims, BoxSize, pos, n, scheme, out, (hipStream_t)stream);
}

