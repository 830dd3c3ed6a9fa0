// Viscous (BR1) FR residual kernels for gfx950, one 64-lane wavefront per cell,
// 4 wavefronts per workgroup sharing the dense operator tables in LDS.
//
//   k_visc_volume:  per cell, from the conserved state U:
//       dQ = primitive(U) - Q_ref, umean = mean_s |U| (outlier reference),
//       grad (u, v, w, T) = J^-T (D field)                  gradients.py:23-90
//       G = viscous flux (tau, heat flux), G~ = adj G       flux_kernels.py:72-137
//       R_vol = sum_m D_m G~_m / det                        viscous_flux.py:113-302
//       h_ip = mean_s(det)^(1/3)                            flux_kernels.py:141-200
//   k_visc_iface: per cell and face slot / flux point, the BR1 common flux of
//       the averaged (Q, grad u, grad T, mu_t) minus the own flux, contracted
//       with the contravariant face row; boundary faces take the ghost state
//       with mirrored gradients and the interior-penalty momentum term
//       (viscous_flux_kernel.py:59-289); + g' distribution / det, outlier
//       zeroing (troubled_cell.py:339-408), R = or += result.
#include <hip/hip_runtime.h>

#include "../../include/afr.h"
#include "afr_physics.hpp"
#include "afr_viscous.hpp"
#include "afr_outlier.hpp"

using namespace afr;

// unroll factors of the interface kernel's own-trace and neighbour gathers (register budget
// against loads in flight)
#ifndef AFR_VISC_TRACE_UNROLL
#define AFR_VISC_TRACE_UNROLL 3
#endif
#ifndef AFR_VISC_GATHER_UNROLL
#define AFR_VISC_GATHER_UNROLL 3
#endif

namespace afr_visc {

constexpr int kWaves = 4;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_fresh(int t) {
    asm volatile("" : "+v"(t));
    return t;
}

template <int P, int CT> struct VO {
    static constexpr int N1 = P + 1, NS = N1 * N1 * N1, NFP = N1 * N1, NSLOT = CT == 0 ? 5 : 4;
    static constexpr int NSV = NS * 5;
};

__device__ __forceinline__ int slot_code(int slot) { return slot < 3 ? slot : (slot == 3 ? 4 : 5); }

// ------------------------------------------------------------------ volume kernel
// Row widths of the sparse collapsed-coordinate derivative matrices: d/da couples the
// N1 points of an a-line, d/db the N1^2 points of an (a,b) plane, d/dc the N1 points of a
// c-line for the prism and all N1^3 points for the tet.
// non-zeros per derivative row: prism (N1, N1^2, N1), tet (N1, N1^2, N1^3); P3 tet rows are
// kept dense (the collapsed-basis round-off entries there are not negligible)
template <int P, int CT> struct DW {
    static constexpr int N1 = P + 1, NS = N1 * N1 * N1;
    static constexpr bool DENSE = CT == 1 && P >= 3;
    static constexpr int W0 = DENSE ? NS : N1, W1 = DENSE ? NS : N1 * N1, W2 = CT == 0 ? N1 : NS;
    static constexpr int WM = W2 > W1 ? W2 : W1;
};

// LDS per wave of the volume kernel: pack [s][kVP] (which also holds the gradients) and
// G~ [m][s][4]; |U| [s][5] and the gradient fields [s][4] live in the G~ space until G~ is
// written.  Occupancy is bounded by this LDS: measured on the 82^3 P2 mesh 2.15 ms with separate
// |U|, field and gradient arrays, 1.91 ms with |U| and the fields in the G~ space (two cells per
// wave instead measured 2.19 ms).
template <int P, int CT> struct VolPlan {
    static constexpr int NS = (P + 1) * (P + 1) * (P + 1);
    static constexpr int PER = NS * kVP + 12 * NS;
};

template <int P, int CT, typename real>
__global__ void __launch_bounds__(64 * kWaves) AFR_WPE_SEL(AFR_WPE_VISC_VOLUME) k_visc_volume(ViscArgs<real> a) {
    using O = VO<P, CT>;
    using W = DW<P, CT>;
    using L = VolPlan<P, CT>;
    constexpr int NS = O::NS, NSV = O::NSV, WM = W::WM;
    static_assert(NSV + 4 * NS <= 12 * NS, "|U| and the gradient fields fit the G~ space");
    // sparse D in LDS unless it is too large (P3 tets: 3 x 64 x 64 entries): then global memory
    constexpr bool D_LDS = 3 * NS * WM * (sizeof(real) + 4) <= 64 * 1024;
    __shared__ real s_Dw_l[D_LDS ? 3 * NS * WM : 1];         // sparse D: weights [m][s][k]
    __shared__ int s_Di_l[D_LDS ? 3 * NS * WM : 1];          //           indices
    const real* const s_Dw = D_LDS ? s_Dw_l : a.Dsp_w;
    const int* const s_Di = D_LDS ? s_Di_l : a.Dsp_i;
    __shared__ __attribute__((aligned(16))) real s_w[kWaves][L::PER];
    __shared__ double s_ref[5];
    const int tid = threadIdx.x, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    int t = tid & 63;
    if constexpr (D_LDS) {
        for (int i = tid; i < 3 * NS * WM; i += 64 * kWaves) {
            s_Dw_l[i] = a.Dsp_w[i];
            s_Di_l[i] = a.Dsp_i[i];
        }
    }
    if (tid < 5) s_ref[tid] = a.qref[tid];
    __syncthreads();
    real* const sPk = s_w[wid];               // interface pack [s][kVP], written out coalesced
    real* const sGt = sPk + NS * kVP;         // G~ [m][s][4] (v = 1..4)
    real* const sU = sGt;                     // |U| [s][5]            (until G~ is written)
    real* const sF = sGt + NSV;               // (du, dv, dw, dT) [s][4] (until G~ is written)
    const double q0r = s_ref[0], q0p = s_ref[4];
    const double T0 = q0p / (rmax(q0r, 1e-10) * kRair);
    constexpr int WD[3] = {W::W0, W::W1, W::W2};

    for (CellWalk cw(a.n_cells, kWaves, wid); cw.it < cw.hi; cw.it += cw.step) {
        const int64_t cell = a.cell_begin + walk_cell(a.schedule, cw.it);
        t = lane_fresh(t);
        // 1. primitives, perturbations, |U|; the lane's adjugate (used in stage 3 by the same
        //    lane) is loaded here with the state
        real mut_s = real(0);
        real adj_s[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) adj_s[q] = real(0);
        if (t < NS) {
            const int s = t;
#pragma unroll
            for (int q = 0; q < 9; ++q) adj_s[q] = a.adj[(cell * NS + s) * 9 + q];
            double u[5], q[5];
            load_state(a.U, cell * NS + s, a.so, cell, u);
            cons2prim(u, q);
            real* vp = sPk + s * kVP;
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const real dq = (real)(q[v] - s_ref[v]);
                a.dQ[(cell * NS + s) * 5 + v] = dq;
                vp[v] = dq;
                sU[s * 5 + v] = (real)fabs(u[v]);
            }
            mut_s = a.mu_t ? a.mu_t[cell * NS + s] : real(0);
            vp[17] = mut_s;
            vp[18] = real(0);
            vp[19] = real(0);
            const double T = q[4] / (rmax(q[0], 1e-10) * kRair);
            sF[s * 4 + 0] = (real)(q[1] - s_ref[1]);
            sF[s * 4 + 1] = (real)(q[2] - s_ref[2]);
            sF[s * 4 + 2] = (real)(q[3] - s_ref[3]);
            sF[s * 4 + 3] = (real)(T - T0);
        }
        wave_sync();
        t = lane_fresh(t);
        if (t < 5) {
            real acc = 0;
            for (int s = 0; s < NS; ++s) acc += sU[s * 5 + t];
            a.umean[cell * 5 + t] = acc / real(NS);
        } else if (t == 5) {
            real acc = 0;
            for (int s = 0; s < NS; ++s) acc += a.det[cell * NS + s];
            a.h_ip[cell] = (real)cbrt((double)(acc / real(NS)));
        }
        // 2. gradients: lane (s, field), sparse derivative rows (fixed trip count, unrolled: the
        //    metric loads of every pass are issued together)
#pragma unroll
        for (int jo = 0; jo < (4 * NS + 63) / 64; ++jo) {
            const int o = t + 64 * jo;
            if (o >= 4 * NS) break;
            const int s = o >> 2, f = o & 3;
            real d[3];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                real acc = 0;
#pragma unroll
                for (int k = 0; k < WD[m]; ++k) {
                    const int idx = (m * NS + s) * WM + k;
                    acc += s_Dw[idx] * sF[s_Di[idx] * 4 + f];
                }
                d[m] = acc;
            }
            const real* ad = a.adj + (cell * NS + s) * 9;
            const real idet = recip(a.det[cell * NS + s]);
            real* vp = sPk + s * kVP;
#pragma unroll
            for (int n = 0; n < 3; ++n)
                vp[5 + f * 3 + n] = (ad[0 * 3 + n] * d[0] + ad[1 * 3 + n] * d[1] + ad[2 * 3 + n] * d[2]) * idet;
        }
        wave_sync();
        t = lane_fresh(t);
        // the cell's pack, contiguous in HBM: whole-line stores
        if constexpr (sizeof(real) == 4 && (NS * kVP) % 4 == 0) {
            float4* dst = reinterpret_cast<float4*>(a.vpack + cell * NS * kVP);
            const float4* src = reinterpret_cast<const float4*>(sPk);
            for (int o = t; o < NS * kVP / 4; o += 64) dst[o] = src[o];
        } else {
            for (int o = t; o < NS * kVP; o += 64) a.vpack[cell * NS * kVP + o] = sPk[o];
        }
        // 3. viscous flux at the solution points, contravariant components (overwrites |U| and
        //    the gradient fields: both are read above, before the last wave_sync)
        if (t < NS) {
            const int s = t;
            real q[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) q[v] = (real)(s_ref[v] + (double)sPk[s * kVP + v]);
            real gv[3][3], gT[3], G[3][5];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
#pragma unroll
                for (int j = 0; j < 3; ++j) gv[i][j] = sPk[s * kVP + 5 + i * 3 + j];
                gT[i] = sPk[s * kVP + 14 + i];
            }
            viscous_flux(q, gv, gT, (real)a.mu, (real)a.Pr, mut_s, (real)a.Prt, G);
            const real* ad = adj_s;
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
                for (int v = 1; v < 5; ++v)
                    sGt[(m * NS + s) * 4 + v - 1] = ad[m * 3 + 0] * G[0][v] + ad[m * 3 + 1] * G[1][v] + ad[m * 3 + 2] * G[2][v];
        }
        wave_sync();
        t = lane_fresh(t);
        // 4. volume term R_vol = sum_m D_m G~_m / det
#pragma unroll
        for (int jo = 0; jo < (NSV + 63) / 64; ++jo) {
            const int o = t + 64 * jo;
            if (o >= NSV) break;
            const int s = o / 5, v = o % 5;
            real acc = 0;
            if (v > 0) {
#pragma unroll
                for (int m = 0; m < 3; ++m)
#pragma unroll
                    for (int k = 0; k < WD[m]; ++k) {
                        const int idx = (m * NS + s) * WM + k;
                        acc += s_Dw[idx] * sGt[(m * NS + s_Di[idx]) * 4 + v - 1];
                    }
                acc *= recip(a.det[cell * NS + s]);
            }
            a.rvol[cell * NSV + o] = acc;
        }
        wave_sync();
    }
}

// ------------------------------------------------------------------ interface kernel
// BR1 jump of one slot / flux point from the own trace po and the other side's trace pn (18
// values each: dQ, grad u, grad T, mu_t); boundary slots take the ghost state, mirrored
// gradients and mu_t, and the interior-penalty momentum term.
template <typename real>
__device__ __forceinline__ void br1_jump(const ViscArgs<real>& a, const double* s_ref, int64_t cell, int64_t sidx,
                                         int slot, int fp, int nfp, int bc, const real* po, real* pn, real mu, real Pr,
                                         real Prt, real* jmp) {
    constexpr int NP = 18;
    const real side = (slot_code(slot) & 1) ? real(1) : real(-1);
    real Qo[5], Qn[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) Qo[v] = (real)(s_ref[v] + (double)po[v]);
    if (bc == BC_INTERIOR) {
#pragma unroll
        for (int v = 0; v < 5; ++v) Qn[v] = (real)(s_ref[v] + (double)pn[v]);
    } else {
        // ghost state (FP64, as in the inviscid path), gradients and mu_t mirrored
        double qod[5], qgd[5];
        const double tn[3] = {(double)a.slot_tn[sidx * 3 + 0], (double)a.slot_tn[sidx * 3 + 1],
                              (double)a.slot_tn[sidx * 3 + 2]};
#pragma unroll
        for (int v = 0; v < 5; ++v) qod[v] = s_ref[v] + (double)po[v];
        ghost_state(bc, qod, tn, a.bcp + (a.cell_case ? 40 * a.cell_case[cell] : 0), qgd);
#pragma unroll
        for (int v = 0; v < 5; ++v) Qn[v] = (real)qgd[v];
#pragma unroll
        for (int k = 0; k < NP; ++k) pn[k] = po[k];
    }
    real Qa[5], gva[3][3], gTa[3], gvo[3][3], gTo[3];
#pragma unroll
    for (int v = 0; v < 5; ++v) Qa[v] = real(0.5) * (Qo[v] + Qn[v]);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            gvo[i][j] = po[5 + i * 3 + j];
            gva[i][j] = real(0.5) * (po[5 + i * 3 + j] + pn[5 + i * 3 + j]);
        }
        gTo[i] = po[14 + i];
        gTa[i] = real(0.5) * (po[14 + i] + pn[14 + i]);
    }
    const real mto = po[17], mta = real(0.5) * (po[17] + pn[17]);
    const real* ar = a.slot_adjrow + (sidx * nfp + fp) * 3;
    const real A0 = ar[0], A1 = ar[1], A2 = ar[2];
    const real Av[3] = {A0, A1, A2};
    // A . (G(common) - G(own)) from the contracted fluxes
    real fc[5], fo[5];
    viscous_flux_dot(Qa, gva, gTa, mu, Pr, mta, Prt, Av, fc);
    viscous_flux_dot(Qo, gvo, gTo, mu, Pr, mto, Prt, Av, fo);
#pragma unroll
    for (int v = 0; v < 5; ++v) jmp[v] = fc[v] - fo[v];
    if (bc != BC_INTERIOR) {
        const real mag = sqrt(A0 * A0 + A1 * A1 + A2 * A2);
        real h = a.h_ip[cell];
        h = h > Tiny<real>::v ? h : Tiny<real>::v;
        const real scale = real(4.0) * (mu + mto) / h * mag * side;
#pragma unroll
        for (int v = 1; v < 4; ++v) jmp[v] -= scale * (Qo[v] - Qn[v]);
    }
}

// acc[q] += w x[q], q < 18, for one pack point (16-byte aligned, global memory or LDS)
template <typename real>
__device__ __forceinline__ void add_pack_point(real* acc, real w, const real* x) {
    if constexpr (sizeof(real) == 4) {
        const float4* x4 = reinterpret_cast<const float4*>(x);
        real xv[20];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const float4 f = x4[c];
            xv[4 * c + 0] = f.x;
            xv[4 * c + 1] = f.y;
            xv[4 * c + 2] = f.z;
            xv[4 * c + 3] = f.w;
        }
#pragma unroll
        for (int q = 0; q < 18; ++q) acc[q] += w * xv[q];
    } else {
#pragma unroll
        for (int q = 0; q < 18; ++q) acc[q] += w * x[q];
    }
}

// Balanced lanes (prisms, P1/P2): a prism's own face-trace rows have N1 non-zeros except on its
// b face (slot 2, N1^2), and of its neighbour rows at most one slot is N1^2 wide on conforming
// prism meshes.  One lane per (slot, flux point) would run every row to the widest row of the
// wave; instead the main lane of (slot, fp) sums the first N1 terms of its rows and (N1 - 1) x NFP
// helper lanes sum the further N1-term parts of the wide own row and of one wide neighbour slot,
// returned through LDS.  Other wide rows (a second wide slot, second-source points) are finished
// by their main lane.
template <int P, int CT, int KMAX, typename real> struct IfaceBal {
    static constexpr int N1 = P + 1, NFP = N1 * N1, NSLOT = CT == 0 ? 5 : 4;
    static constexpr int NM = NSLOT * NFP, NX = (N1 - 1) * NFP;
    static constexpr bool ON = CT == 0 && KMAX == N1 * N1 && NM + NX <= 64;
};

template <int P, int CT, int KMAX, typename real>
__global__ void __launch_bounds__(64 * kWaves) AFR_WPE_SEL(AFR_WPE_VISC_IFACE) k_visc_iface(ViscArgs<real> a) {
    using O = VO<P, CT>;
    using B = IfaceBal<P, CT, KMAX, real>;
    constexpr int N1 = O::N1, NS = O::NS, NFP = O::NFP, NSLOT = O::NSLOT, NSV = O::NSV;
    constexpr int KE = CT == 0 ? N1 * N1 : NS;  // non-zeros of a face-trace row
    constexpr int NP = 18;                      // pack: dQ (5), grad u (9), grad T (3), mu_t (1)
    __shared__ real s_Ew[5 * NFP * KE];
    __shared__ int s_Ei[5 * NFP * KE];
    __shared__ real s_g[2 * N1];
    __shared__ int s_fp[3 * NS], s_co[3 * NS];
    __shared__ __attribute__((aligned(16))) real s_w[kWaves][NS * kVP + 5 * NFP * 5 + NSV];
    __shared__ double s_ref[5];
    // balanced lanes: the neighbour interpolation rows in LDS when the mesh has few row blocks
    // (deduplicated, e.g. 8 blocks for a graded prism box; a batch of cases has more and reads
    // them from global memory)
    constexpr int ROWS_CAP = B::ON ? 8 * NFP + 1 : 1;
    __shared__ real s_rw[ROWS_CAP * KMAX];
    __shared__ int s_ri[ROWS_CAP * KMAX];
    const bool rows_lds = B::ON && a.n_rows + 1 <= ROWS_CAP;
    const int tid = threadIdx.x, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    int t = tid & 63;
    for (int i = tid; i < 5 * NFP * KE; i += 64 * kWaves) {
        s_Ew[i] = a.Esp_w[i];
        s_Ei[i] = a.Esp_i[i];
    }
    if (rows_lds) {
        for (int i = tid; i < (a.n_rows + 1) * KMAX; i += 64 * kWaves) {
            s_rw[i] = a.rows_w[i];
            s_ri[i] = a.rows_idx[i];
        }
    }
    for (int i = tid; i < 3 * NS; i += 64 * kWaves) {
        s_fp[i] = a.fp_of_sp[i];
        s_co[i] = a.coord_of_sp[i];
    }
    if (tid < N1) {
        s_g[tid] = a.gL[tid];
        s_g[N1 + tid] = a.gR[tid];
    }
    if (tid < 5) s_ref[tid] = a.qref[tid];
    __syncthreads();
    real* const sP = s_w[wid];           // own pack [s][kVP]
    real* const sJ = sP + NS * kVP;      // jumps [slot][fp][5]
    real* const sR = sJ + 5 * NFP * 5;   // residual [s][5]
    const real mu = (real)a.mu, Pr = (real)a.Pr, Prt = (real)a.Prt;

    for (CellWalk cw(a.n_cells, kWaves, wid); cw.it < cw.hi; cw.it += cw.step) {
        const int64_t cell = a.cell_begin + walk_cell(a.schedule, cw.it);
        t = lane_fresh(t);
        // balanced lanes: the main lanes' slot tables are loaded with the own pack (independent
        // addresses, one memory round trip for both)
        int m_bc = BC_INTERIOR, m_mask = 0, m_blk0 = -1, m_blk1 = -1, m_nb0 = -1, m_nb1 = -1;
        if constexpr (B::ON) {
            if (t < B::NM) {
                const int64_t sidx = cell * 5 + t / NFP;
                m_bc = a.slot_bc[sidx];
                m_mask = a.slot_mask[sidx];
                m_blk0 = a.slot_blk[sidx * 2];
                m_blk1 = a.slot_blk[sidx * 2 + 1];
                m_nb0 = a.slot_nb[sidx * 2];
                m_nb1 = a.slot_nb[sidx * 2 + 1];
            }
        }
        if constexpr (sizeof(real) == 4 && (NS * kVP) % 4 == 0 &&
                      (NS * kVP + 5 * NFP * 5 + NSV) % 4 == 0) {   // 16-byte aligned rows and waves
            const float4* src = reinterpret_cast<const float4*>(a.vpack + cell * NS * kVP);
            float4* dst = reinterpret_cast<float4*>(sP);
            for (int o = t; o < NS * kVP / 4; o += 64) dst[o] = src[o];
        } else {
            for (int o = t; o < NS * kVP; o += 64) sP[o] = a.vpack[cell * NS * kVP + o];
        }
        wave_sync();
        t = lane_fresh(t);
        // 1. BR1 jumps: lane = (slot, flux point)
        if constexpr (B::ON) {
            constexpr int NM = B::NM, NX = B::NX, OWN_WIDE = 2;
            static_assert((N1 - 1) * NFP * 2 * NP <= NS * kVP + 5 * NFP * 5, "helper traces fit the pack + jump space");
            const bool is_main = t < NM, is_help = t >= NM && t < NM + NX;
            const int fp = is_main ? t % NFP : (t - NM) % NFP;
            const int part = is_main ? 0 : 1 + (t - NM) / NFP;
            const int slot = is_main ? t / NFP : OWN_WIDE;     // own row of this lane
            const int64_t sidx = cell * 5 + slot;
            int bc = BC_INTERIOR, nrow = a.n_rows;
            int64_t nnb = 0;
            bool live = false;
            if (is_main) {
                bc = m_bc;
                if (bc == BC_INTERIOR) {
                    const int src = (m_mask >> fp) & 1;
                    const int row = (src ? m_blk1 : m_blk0) * NFP + fp;
                    const int nbc = src ? m_nb1 : m_nb0;
                    live = nbc >= 0 && row >= 0 && row < a.n_rows;
                    if (live) {
                        nrow = row;
                        nnb = nbc;
                    }
                    live = live && src == 0;    // helper parts only for first-source rows
                }
            }
            // the first main lane with a wide first-source neighbour row names the helped slot
            const bool wide = live && ((m_mask >> 30) & 1);   // host-set: first-source rows wider than N1
            const unsigned long long wm = __ballot(wide);
            const int wslot = wm ? (int)((__ffsll((long long)wm) - 1) / NFP) : -1;
            const bool use_help = is_main && live && slot == wslot;
            // helpers take the helped slot's row and source cell from its main lane of the same
            // flux point (cross-lane, no memory round trip)
            {
                const int from = (wslot >= 0 ? wslot : 0) * NFP + fp;
                const int h_live = __shfl(live ? 1 : 0, from);
                const int h_row = __shfl(nrow, from);
                const int h_nb = __shfl((int)nnb, from);
                if (is_help && wslot >= 0 && h_live) {
                    live = true;
                    nrow = h_row;
                    nnb = h_nb;
                }
            }
            const bool has_nb = is_main ? (bc == BC_INTERIOR && nrow < a.n_rows) : live;
            real po[NP], pn[NP];
#pragma unroll
            for (int k = 0; k < NP; ++k) po[k] = pn[k] = real(0);
            if (is_main || is_help) {
                const int eb = (slot * NFP + fp) * KE + part * N1;
#pragma unroll
                for (int j = 0; j < N1; ++j) {
                    const real w = s_Ew[eb + j];
                    const real* x = sP + s_Ei[eb + j] * kVP;
#pragma unroll
                    for (int q = 0; q < NP; ++q) po[q] += w * x[q];
                }
                if (is_main && slot != OWN_WIDE) {
#pragma unroll 1
                    for (int k = N1; k < KE; ++k) {   // rows are non-zeros first
                        const real w = s_Ew[eb + k];
                        if (w == real(0)) break;
                        const real* x = sP + s_Ei[eb + k] * kVP;
#pragma unroll
                        for (int q = 0; q < NP; ++q) po[q] += w * x[q];
                    }
                }
                if (has_nb) {
                    const int* const ri = a.rows_idx + (int64_t)nrow * KMAX;
                    const real* const rw = a.rows_w + (int64_t)nrow * KMAX;
                    const int k0 = part * N1;
                    int ik[N1];
                    real wk[N1];
                    if (rows_lds) {
#pragma unroll
                        for (int j = 0; j < N1; ++j) {
                            ik[j] = s_ri[nrow * KMAX + k0 + j];
                            wk[j] = s_rw[nrow * KMAX + k0 + j];
                        }
                    } else {
#pragma unroll
                        for (int j = 0; j < N1; ++j) {
                            ik[j] = ri[k0 + j];
                            wk[j] = rw[k0 + j];
                        }
                    }
#pragma unroll
                    for (int j = 0; j < N1; ++j) add_pack_point(pn, wk[j], a.vpack + (nnb * NS + ik[j]) * kVP);
                    if (is_main && !use_help) {
#pragma unroll 1
                        for (int k = N1; k < KMAX; ++k) {
                            const real w = rw[k];
                            if (w == real(0)) break;
                            add_pack_point(pn, w, a.vpack + (nnb * NS + ri[k]) * kVP);
                        }
                    }
                }
            }
            wave_sync();
            // helper parts through LDS (pack and jump space, free until the jumps are written)
            real* const scr = sP;
            if (is_help) {
                real* d = scr + ((part - 1) * NFP + fp) * 2 * NP;
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    d[q] = po[q];
                    d[NP + q] = pn[q];
                }
            }
            wave_sync();
            if (is_main) {
                if (slot == OWN_WIDE) {
#pragma unroll
                    for (int pp = 1; pp < N1; ++pp)
#pragma unroll
                        for (int q = 0; q < NP; ++q) po[q] += scr[((pp - 1) * NFP + fp) * 2 * NP + q];
                }
                if (use_help) {
#pragma unroll
                    for (int pp = 1; pp < N1; ++pp)
#pragma unroll
                        for (int q = 0; q < NP; ++q) pn[q] += scr[((pp - 1) * NFP + fp) * 2 * NP + NP + q];
                }
            }
            wave_sync();
            if (is_main) {
                real jmp[5];
                br1_jump(a, s_ref, cell, sidx, slot, fp, NFP, bc, po, pn, mu, Pr, Prt, jmp);
#pragma unroll
                for (int v = 0; v < 5; ++v) sJ[(slot * NFP + fp) * 5 + v] = jmp[v];
            }
        } else if (t < NSLOT * NFP) {
            const int slot = t / NFP, fp = t % NFP;
            const int64_t sidx = cell * 5 + slot;
            const int bc = a.slot_bc[sidx];
            real po[NP], pn[NP];
#pragma unroll
            for (int k = 0; k < NP; ++k) po[k] = real(0);
#pragma unroll AFR_VISC_TRACE_UNROLL
            for (int k = 0; k < KE; ++k) {
                const int e = (slot * NFP + fp) * KE + k;
                const real w = s_Ew[e];
                const real* x = sP + s_Ei[e] * kVP;
#pragma unroll
                for (int q = 0; q < NP; ++q) po[q] += w * x[q];
            }
            if (bc == BC_INTERIOR) {
#pragma unroll
                for (int k = 0; k < NP; ++k) pn[k] = real(0);
                const int src = (a.slot_mask[sidx] >> fp) & 1;
                int row = a.slot_blk[sidx * 2 + src] * NFP + fp;
                const int nbc = a.slot_nb[sidx * 2 + src];
                if (row < 0 || row >= a.n_rows || nbc < 0) row = a.n_rows;  // all-zero row
                const int64_t nb = nbc >= 0 ? nbc : 0;
                // rows are KMAX wide (zero-padded); three neighbour points in flight at a time:
                // enough loads to cover the latency without the register cost (and lost
                // occupancy) of all KMAX at once
                const int* const ri = a.rows_idx + (int64_t)row * KMAX;
                const real* const rwp = a.rows_w + (int64_t)row * KMAX;
#pragma unroll AFR_VISC_GATHER_UNROLL
                for (int k = 0; k < KMAX; ++k) add_pack_point(pn, rwp[k], a.vpack + (nb * NS + ri[k]) * kVP);
            }
            real jmp[5];
            br1_jump(a, s_ref, cell, sidx, slot, fp, NFP, bc, po, pn, mu, Pr, Prt, jmp);
#pragma unroll
            for (int v = 0; v < 5; ++v) sJ[(slot * NFP + fp) * 5 + v] = jmp[v];
        }
        wave_sync();
        t = lane_fresh(t);
        // 2. distribution, / det, + volume term; the accumulated R is read here too, so that its
        //    latency overlaps the outlier phase instead of ending the cell
        constexpr int NO = (NSV + 63) / 64;
        real r_old[NO];
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            const int o = t + 64 * j;
            r_old[j] = real(0);
            if (o < NSV) {
                const int s = o / 5, v = o % 5;
                if (a.accumulate) r_old[j] = a.R[cell * NSV + o];
                real corr = 0;
#pragma unroll
                for (int slot = 0; slot < NSLOT; ++slot) {
                    const int code = slot_code(slot), axis = code >> 1;
                    const real g = s_g[((code & 1) ? N1 : 0) + s_co[axis * NS + s]];
                    corr += g * sJ[(slot * NFP + s_fp[axis * NS + s]) * 5 + v];
                }
                sR[o] = a.rvol[cell * NSV + o] + corr * recip(a.det[cell * NS + s]);
            }
        }
        wave_sync();
        t = lane_fresh(t);
        if (a.suppress) {
            const real* um = a.umean + cell * 5;
            outlier_zeroing<NS, 5, real>(sR, [um](int v) { return um[v]; }, t);
        }
        wave_sync();
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            const int o = t + 64 * j;
            if (o < NSV) a.R[cell * NSV + o] = a.accumulate ? r_old[j] + sR[o] : sR[o];
        }
        wave_sync();
    }
}

template <int P, int CT, typename real>
int launch_viscous_phase(const ViscArgs<real>& a, int phase, hipStream_t st) {
    if (a.n_cells <= 0) return 0;
    const int grid = walk_grid(a.n_cells, kWaves, a.grid_cap > 0 ? a.grid_cap : 2048);
    if (phase == 0) {
        hipLaunchKernelGGL((k_visc_volume<P, CT, real>), dim3(grid), dim3(64 * kWaves), 0, st, a);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    constexpr int NS = VO<P, CT>::NS, KS = (P + 1) * (P + 1);
    if (a.knz == KS)
        hipLaunchKernelGGL((k_visc_iface<P, CT, KS, real>), dim3(grid), dim3(64 * kWaves), 0, st, a);
    else if (a.knz == NS)
        hipLaunchKernelGGL((k_visc_iface<P, CT, NS, real>), dim3(grid), dim3(64 * kWaves), 0, st, a);
    else
        return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

#define AFR_VISC_INST(P, CT, T) template int launch_viscous_phase<P, CT, T>(const ViscArgs<T>&, int, hipStream_t);
AFR_VISC_INST(1, 0, float)
AFR_VISC_INST(2, 0, float)
AFR_VISC_INST(1, 1, float)
AFR_VISC_INST(2, 1, float)
AFR_VISC_INST(1, 0, double)
AFR_VISC_INST(2, 0, double)
AFR_VISC_INST(1, 1, double)
AFR_VISC_INST(2, 1, double)
AFR_VISC_INST(3, 1, float)
AFR_VISC_INST(3, 1, double)

}  // namespace afr_visc
