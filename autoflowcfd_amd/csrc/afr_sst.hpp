// Shared declarations of the SST k-omega kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "afr_physics.hpp"

namespace afr_sst {

constexpr int kPack = 8;  // face pack: rho, u, v, w, k, omega, Gamma_k, Gamma_omega
constexpr int kInc = 8;   // face-record incidences per cell (prism: 3 quads x 2 + 2 triangles)

template <typename real>
struct SstArgs {
    int64_t cell_begin, n_cells;  // cells of one type
    int64_t n_records;
    const real* U;                // (C, NS, 5) (offset from so's table when it has one)
    afr::StateOff so;
    real* k;                      // (C, NS) in/out
    real* w;                      // (C, NS) in/out
    const real* d_wall;           // (C, NS)
    const real* dt;               // (C, NS)
    real* nu_t;                   // (C, NS) out
    real* mu_t;                   // (C, NS) out: rho(U) nu_t
    real* pack;                   // (C, 5, NFP, kPack) scratch: owner-side face traces of the pack
    real* nbp;                    // (C, NS, 4) scratch: k, omega, Gamma_k, Gamma_omega at the solution points
    real* rates;                  // (C, NS, 2) out when non-null: dk/dt, domega/dt instead of the update
    real* vol;                    // (C, NS, 2) scratch: volume transport terms of k, omega
    real* src;                    // (C, NS, 3) scratch: Sk/rho, Somega/rho, beta blend
    real* crec;                   // (F, NFP, 2) scratch: record corrections
    const real* det;
    const real* adj;              // (C, NS, 3, 3)
    const real* D;                // (3, NS, NS) of the cell type
    const real* Dsp_w;            // (3, NS, WM) sparse derivative rows of the cell type: weights
    const int* Dsp_i;             //                                                     point indices
    const real* Esp_w;            // (5, NFP, KE) sparse face-trace rows of the cell type (slot order)
    const int* Esp_i;
    const real* E0;               // (5, NFP, NS) prism, slot order
    const real* E1;               // (5, NFP, NS) tet, slot order
    const int* slot_nb;
    const int* slot_blk;
    const int* slot_mask;
    const int* rows_idx;
    const real* rows_w;
    int knz, n_rows;
    int64_t n_prism;
    const int* rec_owner;         // (F)
    const int* rec_oslot;         // (F) owner slot index
    const int* rec_flags;         // (F) bit 0: owner-primary interior record (neighbour sources), bit 1: WALL
    const real* rec_tn;           // (F, 3) record unit normal
    const int* inc;               // (C, kInc) record id * 16 + role * 8 + face code  (-1: none)
    const real* gL;
    const real* gR;
    const int* fp_of_sp;
    const int* coord_of_sp;
    double mu_molecular;
    const int* schedule;          // null, or the locality order of the cells of this type
    int grid_cap = 0;     // persistent grid size (workgroups); 0: built-in default
    int grid_cap_face = 0;  // same for the face-record kernel
    int variant = 0;        // cell kernel: 0 two prisms per wavefront, 1 one cell per wavefront
};

template <int P, int CT, typename real>
int launch_sst_cell(const SstArgs<real>& a, hipStream_t st);
template <int P, typename real>
int launch_sst_face(const SstArgs<real>& a, hipStream_t st);
template <int P, int CT, typename real>
int launch_sst_update(const SstArgs<real>& a, hipStream_t st);

int launch_wall_distance(const double* sp_xyz, int64_t n_pts, const double* wall_xyz, int n_wall, void* d_out,
                         int dtype, hipStream_t st);

}  // namespace afr_sst
