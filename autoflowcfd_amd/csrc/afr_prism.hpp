// Shared declarations of the prism (Kronecker-factored) kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "afr_physics.hpp"

struct Ref5 {
    double q[5];
};

namespace afr_prism {

template <int P> struct PK {
    static constexpr int N1 = P + 1;
    static constexpr int NAB = N1 * N1;
    static constexpr int OV = (2 * P < 3 ? 2 * P : 3);
    static constexpr int NF1 = OV + 1;
    static constexpr int NABF = NF1 * NF1;
    static constexpr int NF = NABF * NF1;
    static constexpr int NS = NAB * N1;
    static constexpr int NFP = N1 * N1;
    // packed factor table (autoflowcfd_amd/fr_tables.py PrismFactors.packed order)
    static constexpr int O_C2FAB = 0;
    static constexpr int O_C2FC = O_C2FAB + NABF * NAB;
    static constexpr int O_AA = O_C2FC + NF1 * N1;
    static constexpr int O_AB = O_AA + NAB * NABF;
    static constexpr int O_BC = O_AB + NAB * NABF;
    static constexpr int O_F2CC = O_BC + N1 * NF1;
    static constexpr int O_F2CAB = O_F2CC + N1 * NF1;
    static constexpr int O_EAB = O_F2CAB + NAB * NABF;
    static constexpr int O_EC = O_EAB + 3 * N1 * NAB;
    static constexpr int O_DAB = O_EC + 2 * N1;
    static constexpr int O_DC = O_DAB + 2 * NAB * NAB;
    static constexpr int TABN = O_DC + N1 * N1;
};

template <typename real>
struct PrismArgs {
    int64_t n_cells;  // prisms occupy cells [0, n_cells)
    const int* schedule;  // (n_cells) processing order of the prisms (locality order)
    int grid_cap = 0;     // persistent grid size (workgroups); 0: built-in default
    int variant = 0;      // 0: own cell staged in LDS; 1: own cell and slot sources staged
    const real* dQ;
    const real* umean;
    real* R;
    const real* det;
    const real* adjf;
    const real* gcl;
    const int* slot_nb;
    const int* slot_blk;
    const int* slot_mask;
    const int* slot_bc;
    const int* slot_fb;
    const real* slot_tn;
    const real* slot_adjrow;
    const int* rows_nnz;
    const int* rows_idx;
    const real* rows_w;
    int knz;
    int n_rows;
    const real* tab;
    const real* gL;
    const real* gR;
    const int* fp_of_sp;
    const int* coord_of_sp;
    const double* bcp;
    const int* cell_case;  // null: one case
    double qref[5];
    double fref[15];
    int suppress;
};

template <int P, typename real>
int launch_prism_inviscid(const PrismArgs<real>& a, hipStream_t st);

template <int NS, typename real>
int launch_prim_umean(const real* U, real* dQ, real* umean, int64_t n, const double* ref, afr::StateOff so,
                      hipStream_t st);

}  // namespace afr_prism
