// Shared declarations of the wall force kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "afr_physics.hpp"

namespace afr_forces {

template <typename real>
struct ForceArgs {
    int64_t n_wall, n_prism;
    const real* U;          // offset from so's table when it has one
    afr::StateOff so;
    const real* mu_t;       // may be null
    const real* det;
    const real* adj;
    const real* D0;         // prism (3, NS, NS)
    const real* D1;         // tet
    const real* E0;         // prism (5, NFP, NS) slot order
    const real* E1;         // tet
    const int* wrec_owner;  // (W)
    const int* wrec_oslot;  // (W)
    const double* wrec_geom;  // (W, 4): unit normal, record area
    const double* w1d;      // (order+1) Gauss-Legendre weights
    double mu;
    double* part;           // (W, 6) scratch
    double* out;            // (6): pressure force, viscous force
};

template <int P, typename real>
int launch_wall_forces(const ForceArgs<real>& a, hipStream_t st);

// per-record force partials only (a.part), and the fixed-order sum of n records of part into out[6]
template <int P, typename real>
int launch_wall_forces_records(const ForceArgs<real>& a, hipStream_t st);
int launch_sum_forces(const double* part, int64_t n, double* out, hipStream_t st);

}  // namespace afr_forces
