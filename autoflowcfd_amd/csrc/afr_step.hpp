// Shared declarations of the pseudo-step kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "afr_physics.hpp"

namespace afr_step {

constexpr int kNormParts = 4096;  // capacity of the FP64 partial-sum buffer

template <typename real>
int launch_dt_local(const real* U, afr::StateOff so, const real* mu_t, const real* det, const real* mfs,
                    const real* vol, real* dt, int64_t n_cells, int ns, double mu, int order, int nut_mode,
                    hipStream_t st);

template <typename real>
int launch_rk_stage(int ns, const real* U0, const real* Ua, const real* R, const real* dt, real* U_out,
                    afr::StateOff so, const real* filt0, const real* filt1, int64_t n_cells, int64_t n_prism,
                    double c0, double c1, double c2, double p_floor, hipStream_t st);

template <typename real>
int launch_norm(const real* x, int64_t n, double n_total, double* part, int n_part_max, double* out, hipStream_t st);

// x[0, n) = v (a fill kernel, not a memset: safe inside captured graphs)
template <typename real>
int launch_fill(real* x, int64_t n, double v, hipStream_t st);

// norm of the logical array x[a0, a0 + n0) ++ x[a1, a1 + n1) (bit-identical to launch_norm of the
// same values stored contiguously)
template <typename real>
int launch_norm2(const real* x, int64_t a0, int64_t n0, int64_t a1, int64_t n1, double n_total, double* part,
                 int n_part_max, double* out, hipStream_t st);

}  // namespace afr_step
