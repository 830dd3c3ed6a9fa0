// Shared declarations of the viscous (BR1) kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "afr_physics.hpp"

namespace afr_visc {

constexpr int kVP = 20;  // per-point viscous pack: dQ (5), grad u (9), grad T (3), mu_t (1), pad (2)

template <typename real>
struct ViscArgs {
    int64_t cell_begin, n_cells;   // cells [cell_begin, cell_begin + n_cells) of one type
    const real* U;                 // (C, NS, 5) conserved (offset from so's table when it has one)
    afr::StateOff so;
    const real* mu_t;              // (C, NS) or null
    real* dQ;                      // (C, NS, 5) out: primitive - Q_ref
    real* umean;                   // (C, 5) out
    real* grad;                    // (C, NS, 12) out: grad u (9, [i][j] = du_i/dx_j), grad T (3)
    real* rvol;                    // (C, NS, 5) out: volume term / det
    real* h_ip;                    // (C) out: mean(det)^(1/3)
    real* R;                       // (C, NS, 5) out
    const real* det;
    const real* adj;               // (C, NS, 3, 3)
    real* vpack;                   // (C, NS, kVP) out: per-point pack for the interface kernel
    const real* Dsp_w;             // (3, NS, NS) sparse derivative rows of the cell type: weights
    const int* Dsp_i;              //                                            indices
    const real* Esp_w;             // (5, NFP, KE) sparse face-trace rows of the cell type, slot order
    const int* Esp_i;
    const int* slot_nb;
    const int* slot_blk;
    const int* slot_mask;
    const int* slot_bc;
    const real* slot_tn;
    const real* slot_adjrow;
    const int* rows_idx;
    const real* rows_w;
    int knz;
    int n_rows;
    const real* gL;
    const real* gR;
    const int* fp_of_sp;
    const int* coord_of_sp;
    const double* bcp;
    const int* cell_case;          // null: one case
    const int* schedule;           // null, or the locality order of the cells of this type
    int grid_cap = 0;     // persistent grid size (workgroups); 0: built-in default
    double qref[5];
    double mu, Pr, Prt;
    int suppress;
    int accumulate;
};

// phase 0: k_visc_volume; phase 1: k_visc_iface
template <int P, int CT, typename real>
int launch_viscous_phase(const ViscArgs<real>& a, int phase, hipStream_t st);

}  // namespace afr_visc
