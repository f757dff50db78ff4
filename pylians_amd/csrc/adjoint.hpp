// Adjoint (gradient of sum_b w_b P0_b w.r.t. positions) and grid interpolation at positions.
#pragma once
#include "common.hpp"

namespace pyl {

// spec: window-corrected half spectrum as written by the binning pass (keep); replaced in place
// by H_k * post (see adjoint.hip).  n3: independent-mode counts per 3-D bin (device).
template <typename R>
int grad_spectrum(cplx<R>* spec, int N, int64_t P, const double* w, int nb, const unsigned long long* n3,
                  double BoxSize, const double* win, double post, hipStream_t st);

// grad[3p + a] = scale * (N/L) * sum_cells G * W'_a prod W_b  (float32 grid G, (n,3) output)
int gather_grad(const float* G, int N, double BoxSize, const float* pos, int64_t n, int scheme, double scale,
                float* grad, hipStream_t st);

// out[p] = sum_cells G * prod W  (MAS_library.CIC_interp with scheme 2)
int interp(const float* G, int N, double BoxSize, const float* pos, int64_t n, int scheme, float* out,
           hipStream_t st);

}  // namespace pyl
