// Mass assignment (particles -> grid) for gfx950.
#pragma once
#include "common.hpp"

namespace pyl {

enum DepositFlags {
    DEP_GRID_F64 = 1,   // grid is float64 (CICW_d / NGPW_d, MAS_library.pyx:229-266, 338-358)
    DEP_DUP2D = 2,      // 2-D Cython path: every weight deposited STENCIL times (MAS_library.pyx:137-139)
    DEP_RENORM2D = 4,   // ... and the whole grid divided by STENCIL afterwards (MAS_library.pyx:84-110)
    DEP_ZERO = 8,       // zero the grid before depositing (MA itself accumulates)
    DEP_ATOMIC = 16,    // force the direct global-atomic kernel
    DEP_BINNED = 32,    // force the cell-binned LDS kernel
    DEP_SORTED = 64,    // force the cell-binned LDS kernel after a presort by supertile
    DEP_ACC_F32W = 1024,   // binned deposit into a float32 grid: float32 weights, float64 cell sums
    DEP_ACC_F64 = 2048,    // binned deposit: float64 weights and cell sums (every cell correctly rounded)
};

// cell accumulation of the binned deposit into a float32 grid: 0 float64 weights and sums (default),
// 1 float32 weights and float64 sums (the ACC_* of binned.hip)
inline int acc_mode(int flags) { return (flags & DEP_ACC_F32W) && !(flags & DEP_ACC_F64) ? 1 : 0; }

// Redshift-space shift fused into the position load (redshift_space_library.pyx:29-43):
// x[axis] <- x[axis] + vel[axis] * fac in float32, fac = (float)((1 + z) / H), and
// (x + L) mod L when the result leaves [0, L].  vel == nullptr: no shift.
struct Rsd {
    const float* vel = nullptr;   // (n, 3) float32, row-major like the positions
    int axis = 0;
    float fac = 0.f;
    float box = 0.f;
};

__device__ __forceinline__ float rsd_shift(float x, float v, float fac, float box) {
#pragma clang fp contract(off)
    const float dv = v * fac;   // two roundings (no fused multiply-add), as the reference's float arithmetic
    float s = x + dv;
    if (s > box || s < 0.0f) s = fmodf(s + box, box);
    return s;
}

// pos[:, axis] shifted in place by rsd_shift with fac = (1 + z) / H; pos and vel (n, cols) float32.
int pos_redshift_space(float* pos, const float* vel, int64_t n, int cols, int axis, float fac, float box,
                       hipStream_t st);

// pos: (n, axes) float32 row-major device array; W nullable; grid (dims^axes) float32 or float64.
// scheme: 1 NGP, 2 CIC, 3 TSC, 4 PCS.  The cell index and sub-cell fractions are computed in
// float64 from the float32 positions; weights are float32 (float64 grid: float64).
int deposit(const float* pos, const float* W, int64_t n, int axes, int dims, double BoxSize, int scheme,
            int flags, void* grid, hipStream_t st, const Rsd& rsd = Rsd());

// grid /= s (float32 or float64 grid of n cells), as numpy's `number2 /= 2.0` (MAS_library.pyx:89)
int grid_divide(void* grid, int64_t n, bool f64, double s, hipStream_t st);
// out[i] = (float)(rho[i] * n / mass - 1) in float64 (mass = dmass[0] when dmass is non-null).
int delta_round(const double* rho, float* out, int64_t n, const double* dmass, double mass, hipStream_t st);
// delta = rho * n / M - 1 in place on a float64 grid of n cells (M: dmass[0] when given, else mass)
int delta_inplace(double* g, int64_t n, const double* dmass, double mass, hipStream_t st);

// Cell-binned LDS deposit (3-D grids with dims % 16 == 0).  mode: 0 = grid += rho, 1 = grid = rho,
// 2 = grid = rho * N^3 / mass - 1 with mass = dmass[0] when dmass != nullptr, else `mass`.
bool binned_supported(int axes, int dims);
int deposit_binned(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, bool f64,
                   int mode, double mult, void* grid, const double* dmass, double mass, hipStream_t st,
                   const Rsd& rsd = Rsd(), bool sorted = false, int acc = 0);

}  // namespace pyl
