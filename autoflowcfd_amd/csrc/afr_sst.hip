// SST k-omega turbulence step on the device (once per pseudo-step, before the
// mean-flow stages), FP64 arithmetic on storage-precision fields.
//
//   k_sst_cell    per cell: strain rate from the gradient of the conserved momentum,
//                 clipped grad k / grad omega, F1, F2, nu_t (viscosity-ratio cap),
//                 production / destruction / cross diffusion with mu = 1.8e-5
//                 (src/autoflowcfd/core/turbulence/sst.py:21-330,
//                  src/autoflowcfd/core/fr_solver/turbulence.py:178-330);
//                 transport volume terms -D.(adj rho u phi)/det, -D.(adj Gamma grad phi)/det
//                 and the face pack (src/autoflowcfd/core/turbulence/transport.py:122-360)
//   k_sst_face    per face record and flux point: upwind convection and BR1-type
//                 diffusion corrections from the owner side, Neumann ghost on records
//                 without neighbour sources, Dirichlet-zero k on WALL records
//                 (src/autoflowcfd/core/turbulence/transport_kernel.py:103-170)
//   k_sst_update  per cell: record corrections gathered with the owner's / neighbour's
//                 flux-point maps (transport_kernel.py:200-260), / rho, outlier zeroing,
//                 point-implicit source damping and explicit update with the local time
//                 step, floors (src/autoflowcfd/core/turbulence/sst.py:335-479)
//   k_wall_distance  nearest wall node per solution point, FP64 brute force
//                 (src/autoflowcfd/core/fr_solver/turbulence.py:112-121 uses a KD-tree)
#include <hip/hip_runtime.h>

#include "../../include/afr.h"
#include "afr_physics.hpp"
#include "afr_sst.hpp"
#include "afr_sst_point.hpp"
#include "afr_outlier.hpp"

using namespace afr;

namespace afr_sst {

constexpr int kWaves = 4;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int slot_code(int slot) { return slot < 3 ? slot : (slot == 3 ? 4 : 5); }

template <int P, int CT> struct SO {
    static constexpr int N1 = P + 1, NS = N1 * N1 * N1, NFP = N1 * N1, NSLOT = CT == 0 ? 5 : 4;
};

// ------------------------------------------------------------------ per-cell sources and volume terms
// One wave per cell.  The contractions (gradients D x, face traces E x, volume divergence D F)
// run in the storage precision on per-cell perturbations x - x_0 (x_0: the value at the cell's
// first solution point; D rows sum to zero and E rows to one, so D x = D (x - x_0) and
// E x = x_0 + E (x - x_0)), which keeps FP32 contractions as accurate as FP64 ones on FP32
// data.  The SST source terms are FP64 (the blend functions FP32 on the FP32 path).
// non-zeros per sparse derivative row (prism: N1, N1^2, N1; tet: N1, N1^2, N1^3) and per
// face-trace row (prism N1^2, tet N1^3): the widths of the mesh's sparse D / E tables
template <int P, int CT> struct SW {
    static constexpr int N1 = P + 1, NS = N1 * N1 * N1;
    static constexpr int W0 = N1, W1 = N1 * N1, W2 = CT == 0 ? N1 : NS;
    static constexpr int WM = W2 > W1 ? W2 : W1;
    static constexpr int KE = CT == 0 ? N1 * N1 : NS;
};

// cells per wavefront of the cell kernel (its tile variant, PAIR): two half-waves of 32 lanes when
// a cell has at most 32 solution points, so the per-point phases (FP64 sources) keep 54 of 64
// lanes busy at P2 (PAIR 1), or one cell per wavefront (PAIR 0); the autotuner picks per mesh
template <int P, int CT, int PAIR> struct SstCellPlan {
    static constexpr int NS = (P + 1) * (P + 1) * (P + 1);
    static constexpr int CPW = (PAIR && CT == 0 && NS <= 32) ? 2 : 1;   // cells per wave (prisms)
    static constexpr int LW = 64 / CPW;                             // lanes per cell
};

template <int P, int CT, typename real, int PAIR>
__global__ void __launch_bounds__(64 * kWaves) AFR_WPE_SEL(AFR_WPE_SST_CELL) k_sst_cell(SstArgs<real> a) {
    using O = SO<P, CT>;
    using W = SW<P, CT>;
    using L = SstCellPlan<P, CT, PAIR>;
    constexpr int NS = O::NS, WM = W::WM, KE = W::KE, CPW = L::CPW, LW = L::LW;
    constexpr int WD[3] = {W::W0, W::W1, W::W2};
    // per-SP record (12): rho, u, v, w, k, omega, Gk, Gw (the face-trace fields, contiguous),
    // m1, m2, m3 (the momentum gradient fields), pad
    constexpr int NF = 12, G_ = 16;  // gradient record: grad m (9), grad k (3), grad omega (3), pad
    constexpr int NFP = O::NFP, NSLOT = O::NSLOT;
    constexpr int PER_CELL = NS * NF + NS * G_ + NS * 8;
    // sparse derivative and face-trace rows (the D / E non-zeros only)
    __shared__ real s_Dw[3 * NS * WM];
    __shared__ int s_Di[3 * NS * WM];
    __shared__ real s_Ew[5 * NFP * KE];
    __shared__ int s_Ei[5 * NFP * KE];
    __shared__ __attribute__((aligned(16))) real s_w[kWaves][CPW * PER_CELL];
    for (int i = threadIdx.x; i < 3 * NS * WM; i += 64 * kWaves) {
        s_Dw[i] = a.Dsp_w[i];
        s_Di[i] = a.Dsp_i[i];
    }
    for (int i = threadIdx.x; i < 5 * NFP * KE; i += 64 * kWaves) {
        s_Ew[i] = a.Esp_w[i];
        s_Ei[i] = a.Esp_i[i];
    }
    __syncthreads();
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), t = threadIdx.x & 63;
    const int h = t / LW, tl = t % LW;               // cell of the wave, lane within the cell
    real* sv = s_w[wid] + h * PER_CELL;  // [s][NF]
    real* sg = sv + NS * NF;             // [s][G_]
    real* sF = sg + NS * G_;             // [j][f][4]: contravariant convective + diffusive flux of k (f=0), omega (f=1), m = 0..2
    const double mu = a.mu_molecular;
    const int64_t n_units = (a.n_cells + CPW - 1) / CPW;
    for (CellWalk cw(n_units, kWaves, wid); cw.it < cw.hi; cw.it += cw.step) {
        // list positions CPW it + h; a missing second cell leaves its half-wave idle
        const int64_t pos0 = cw.it * CPW;
        const int64_t c0 = walk_cell(a.schedule, pos0);
        const int64_t c1 = (CPW == 2 && pos0 + 1 < a.n_cells) ? walk_cell(a.schedule, pos0 + 1) : c0;
        const bool live = h == 0 || pos0 + 1 < a.n_cells;
        const int64_t cell = a.cell_begin + (h == 0 ? c0 : c1);
        const int64_t cb = cell * NS;                // first solution point of the cell
        // the lane's wall distance and adjugate (used by the same lane in the point phase) are
        // loaded here with the state
        real d_s = real(0), adj_s[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) adj_s[q] = real(0);
        if (live && tl < NS) {
            const int64_t i = cb + tl;
            d_s = a.d_wall[i];
#pragma unroll
            for (int q = 0; q < 9; ++q) adj_s[q] = a.adj[i * 9 + q];
            real u5[5];
            double off[5];
            a.so.row(cell, off);
#pragma unroll
            for (int v = 0; v < 5; ++v)
                u5[v] = a.so.mode ? (real)((double)a.U[i * 5 + v] + off[v]) : a.U[i * 5 + v];
            const real rho = u5[0] > real(1e-10) ? u5[0] : real(1e-10);
            real* r = sv + tl * NF;
            r[0] = rho;
            r[1] = u5[1] / rho;
            r[2] = u5[2] / rho;
            r[3] = u5[3] / rho;
            r[4] = a.k[i];
            r[5] = a.w[i];
            r[8] = u5[1];
            r[9] = u5[2];
            r[10] = u5[3];
        }
        wave_sync();
        // gradients of m1, m2, m3, k, omega: lane (s, field), sparse rows, perturbations about SP 0
        for (int o = tl; o < 5 * NS; o += LW) {
            if (!live) break;
            const int s = o / 5, f = o % 5;
            const int src = f < 3 ? 8 + f : 4 + (f - 3);
            const real x0 = sv[src];
            real d[3];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                real acc = 0;
#pragma unroll
                for (int k = 0; k < WD[m]; ++k) {
                    const int idx = (m * NS + s) * WM + k;
                    acc += s_Dw[idx] * (sv[s_Di[idx] * NF + src] - x0);
                }
                d[m] = acc;
            }
            const real* ad = a.adj + (cb + s) * 9;
            const real idet = real(1) / a.det[cb + s];
            const int base = f < 3 ? f * 3 : 9 + (f - 3) * 3;
#pragma unroll
            for (int n = 0; n < 3; ++n) sg[s * G_ + base + n] = (ad[0 * 3 + n] * d[0] + ad[1 * 3 + n] * d[1] + ad[2 * 3 + n] * d[2]) * idet;
        }
        wave_sync();
        if (live && tl < NS) {
            const int s = tl;
            const int64_t i = cb + s;
            const real* r = sv + s * NF;
            const real* gs = sg + s * G_;
            const double rho = (double)r[0], k = (double)r[4], w = (double)r[5];
            const double d = (double)d_s;
            // strain rate of the momentum gradient
            double S2 = 0.0;
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double sij = 0.5 * ((double)gs[p * 3 + q] + (double)gs[q * 3 + p]);
                    S2 += sij * sij;
                }
            const double S = sqrt(2.0 * S2);
            // clipped gradients of k and omega
            double gk[3], gw[3];
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                gk[n] = (double)gs[9 + n];
                gw[n] = (double)gs[12 + n];
            }
            const double mk = sqrt(gk[0] * gk[0] + gk[1] * gk[1] + gk[2] * gk[2]);
            const double mw = sqrt(gw[0] * gw[0] + gw[1] * gw[1] + gw[2] * gw[2]);
            if (mk > 1e6) {
                const double f = dmin(dmax(1e6 / dmax(mk, 1e-10), 0.0), 1.0);
                gk[0] *= f; gk[1] *= f; gk[2] *= f;
            }
            if (mw > 1e6) {
                const double f = dmin(dmax(1e6 / dmax(mw, 1e-10), 0.0), 1.0);
                gw[0] *= f; gw[1] *= f; gw[2] *= f;
            }
            const double gdot = gk[0] * gw[0] + gk[1] * gw[1] + gk[2] * gw[2];
            // sources (mu = 1.8e-5 as the reference's source terms)
            const double musrc = 1.8e-5;
            const double irho = recip(dmax(rho, 1e-10));   // one reciprocal for the four uses
            const double nu = musrc * irho;
            const double ks = dmin(k, 1e40), wsf = dmax(dmin(w, 1e40), 1e-10);
            const double CD = dmax(2.0 * rho * SW2 / wsf * gdot, 1e-10);
            const double F1 = Blend<real>::F1(ks, wsf, d, nu, rho, CD);
            const double F2 = Blend<real>::F2(ks, wsf, d, nu);
            const double beta = F1 * B1 + (1.0 - F1) * B2;
            double nut;
            {
                const double kk = dmax(ks, 1e-10), ww = dmax(wsf, 1e-10), SS = dmax(S, 1e-10);
                nut = A1 * kk / dmax(A1 * ww, F2 * SS);
                nut = dmin(nut, 1.0e5 * musrc * irho);
                nut = isfinite(nut) ? nut : 0.0;
            }
            const double Pk = dmin(nut * rho * S * S, 10.0 * BSTAR * rho * ks * wsf);
            double Sk = Pk - rho * BSTAR * ks * wsf;
            const double g1 = B1 / BSTAR - SW1 * KAP * KAP / sqrt(BSTAR);
            const double g2 = B2 / BSTAR - SW2 * KAP * KAP / sqrt(BSTAR);
            const double gam = F1 * g1 + (1.0 - F1) * g2;
            double Sw = rho * gam * S * S - rho * beta * wsf * wsf + 2.0 * rho * (1.0 - F1) * SW2 / wsf * gdot;
            Sk = finite_or(Sk, 0.0);
            Sw = finite_or(Sw, 0.0);
            const double ir = irho;
            a.src[i * 3 + 0] = (real)(Sk * ir);
            a.src[i * 3 + 1] = (real)(Sw * ir);
            a.src[i * 3 + 2] = (real)beta;
            a.nu_t[i] = (real)nut;
            a.mu_t[i] = (real)(rho * nut);
            // diffusivities with the transport's own F1 (molecular mu, raw k / omega)
            const double wsaf = dmax(w, 1e-10);
            const double CDt = dmax(2.0 * rho * SW2 / wsaf * gdot, 1e-10);
            const double F1t = Blend<real>::F1(k, w, d, mu * irho, rho, CDt);
            const double rnut = rho * nut;
            const double Gk = mu + (F1t * SK1 + (1.0 - F1t) * SK2) * rnut;
            const double Gw = mu + (F1t * SW1 + (1.0 - F1t) * SW2) * rnut;
            real* rw = sv + s * NF;
            rw[6] = (real)Gk;
            rw[7] = (real)Gw;
            real* nbp = a.nbp + i * 4;
            nbp[0] = (real)k;
            nbp[1] = (real)w;
            nbp[2] = (real)Gk;
            nbp[3] = (real)Gw;
            // contravariant transport fluxes (raw gradients in the diffusion), convective +
            // diffusive summed per field: sF[s][f][m]
            const real* ad = adj_s;
            const real rr = r[0];
            const real ru[3] = {rr * r[1], rr * r[2], rr * r[3]};
            const real gkr = (real)Gk, gwr = (real)Gw, kr = r[4], wr = r[5];
            real* fo = sF + s * 8;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const real am[3] = {ad[m * 3 + 0], ad[m * 3 + 1], ad[m * 3 + 2]};
                const real fr = am[0] * ru[0] + am[1] * ru[1] + am[2] * ru[2];
                fo[0 * 4 + m] = fr * kr + gkr * (am[0] * gs[9] + am[1] * gs[10] + am[2] * gs[11]);
                fo[1 * 4 + m] = fr * wr + gwr * (am[0] * gs[12] + am[1] * gs[13] + am[2] * gs[14]);
            }
        }
        wave_sync();
        // owner-side face traces of (rho, u, v, w, k, omega, Gamma_k, Gamma_omega): lane (slot, flux point)
        for (int o = tl; o < NSLOT * NFP; o += LW) {
            if (!live) break;
            const int slot = o / NFP, fp = o % NFP;
            real x0[kPack], tr[kPack];
#pragma unroll
            for (int q = 0; q < kPack; ++q) {
                x0[q] = sv[q];
                tr[q] = 0;
            }
#pragma unroll
            for (int k = 0; k < KE; ++k) {
                const int e = (slot * NFP + fp) * KE + k;
                const real ej = s_Ew[e];
                const real* x = sv + s_Ei[e] * NF;
#pragma unroll
                for (int q = 0; q < kPack; ++q) tr[q] += ej * (x[q] - x0[q]);
            }
            real* ov = a.pack + ((cell * 5 + slot) * NFP + fp) * kPack;
#pragma unroll
            for (int q = 0; q < kPack; ++q) ov[q] = x0[q] + tr[q];
        }
        // volume terms: -(D.F)/det for k and omega, perturbations about SP 0
        for (int o = tl; o < 2 * NS; o += LW) {
            if (!live) break;
            const int s = o >> 1, f = o & 1;
            real acc = 0;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const real f0 = sF[f * 4 + m];
#pragma unroll
                for (int k = 0; k < WD[m]; ++k) {
                    const int idx = (m * NS + s) * WM + k;
                    acc += s_Dw[idx] * (sF[s_Di[idx] * 8 + f * 4 + m] - f0);
                }
            }
            a.vol[(cb + s) * 2 + f] = -acc / a.det[cb + s];
        }
        wave_sync();
    }
}

// ------------------------------------------------------------------ per face record corrections
template <int P, int KMAX, typename real>
__global__ void __launch_bounds__(256) k_sst_face(SstArgs<real> a) {
    constexpr int N1 = P + 1, NS = N1 * N1 * N1, NFP = N1 * N1;
    // the deduplicated neighbour row blocks in LDS when they fit (8 blocks for a graded prism box)
    constexpr int ROWS_CAP = 8 * NFP + 1;
    __shared__ real s_rw[ROWS_CAP * KMAX];
    __shared__ int s_ri[ROWS_CAP * KMAX];
    const bool rows_lds = a.n_rows + 1 <= ROWS_CAP;
    if (rows_lds) {
        for (int i = threadIdx.x; i < (a.n_rows + 1) * KMAX; i += 256) {
            s_rw[i] = a.rows_w[i];
            s_ri[i] = a.rows_idx[i];
        }
    }
    __syncthreads();
    const int64_t n = a.n_records * NFP;
    for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < n; g += (int64_t)gridDim.x * 256) {
        const int64_t r = g / NFP;
        const int fp = (int)(g % NFP);
        const int64_t o = a.rec_owner[r];
        const int slot = a.rec_oslot[r];
        const int flags = a.rec_flags[r];
        double po[kPack], pn[4];
        const real* tr = a.pack + ((o * 5 + slot) * NFP + fp) * kPack;
#pragma unroll
        for (int q = 0; q < kPack; ++q) po[q] = (double)tr[q];
        if (flags & 1) {
            // neighbour k, omega, Gammas at the owner's flux points through the owner's face slot
            const int64_t sidx = o * 5 + slot;
            const int src = (a.slot_mask[sidx] >> fp) & 1;
            int row = a.slot_blk[sidx * 2 + src] * NFP + fp;
            const int nbc = a.slot_nb[sidx * 2 + src];
            if (row < 0 || row >= a.n_rows || nbc < 0) row = a.n_rows;
            const int64_t nb = nbc >= 0 ? nbc : 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) pn[q] = 0.0;
            int spk[KMAX];
            real wk[KMAX];
            if (rows_lds) {
#pragma unroll
                for (int kk = 0; kk < KMAX; ++kk) {  // rows are KMAX wide (zero-padded)
                    spk[kk] = s_ri[row * KMAX + kk];
                    wk[kk] = s_rw[row * KMAX + kk];
                }
            } else {
#pragma unroll
                for (int kk = 0; kk < KMAX; ++kk) {
                    spk[kk] = a.rows_idx[(int64_t)row * KMAX + kk];
                    wk[kk] = a.rows_w[(int64_t)row * KMAX + kk];
                }
            }
#pragma unroll
            for (int kk = 0; kk < KMAX; ++kk) {
                const real* pk = a.nbp + (nb * NS + spk[kk]) * 4;
                const double wgt = (double)wk[kk];
#pragma unroll
                for (int q = 0; q < 4; ++q) pn[q] += wgt * (double)pk[q];
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) pn[q] = po[4 + q];
            if (flags & 2) pn[0] = -po[4];  // k = 0 on no-slip walls: mirrored ghost
        }
        const double n0 = (double)a.rec_tn[r * 3 + 0], n1 = (double)a.rec_tn[r * 3 + 1], n2 = (double)a.rec_tn[r * 3 + 2];
        const double mass = po[0] * (po[1] * n0 + po[2] * n1 + po[3] * n2);
        double c[2];
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            const double phio = po[4 + f], phin = pn[f];
            const double up = mass >= 0.0 ? phio : phin;
            const double gface = 0.5 * (po[6 + f] + pn[2 + f]);
            c[f] = mass * (up - phio) - gface * (0.5 * (phio + phin) - phio);
        }
        a.crec[g * 2 + 0] = (real)c[0];
        a.crec[g * 2 + 1] = (real)c[1];
    }
}

// ------------------------------------------------------------------ gather, outlier zeroing, update
template <int P, int CT, typename real>
__global__ void __launch_bounds__(64 * kWaves) AFR_WPE_SEL(AFR_WPE_SST_UPDATE) k_sst_update(SstArgs<real> a) {
    using O = SO<P, CT>;
    constexpr int N1 = O::N1, NS = O::NS, NFP = O::NFP;
    __shared__ double s_t[kWaves][2 * NS];
    __shared__ double s_fld[kWaves][2 * NS];
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), t = threadIdx.x & 63;
    // per-point tables of the three cube axes (flux point of the point on each axis and the
    // correction-function values at its coordinate), the same for every cell: staged once
    __shared__ int s_fps[3 * NS];
    __shared__ real s_gl[3 * NS], s_gr[3 * NS];
    for (int q = threadIdx.x; q < 3 * NS; q += 64 * kWaves) {
        s_fps[q] = a.fp_of_sp[q];
        const int cs = a.coord_of_sp[q];
        s_gl[q] = a.gL[cs];
        s_gr[q] = a.gR[cs];
    }
    __syncthreads();
    double* st = s_t[wid];
    double* sf = s_fld[wid];
    for (CellWalk cw(a.n_cells, kWaves, wid); cw.it < cw.hi; cw.it += cw.step) {
        const int64_t cell = a.cell_begin + walk_cell(a.schedule, cw.it);
        if (t < 2 * NS) {
            const int s = t >> 1, f = t & 1;
            const int64_t i = cell * NS + s;
            const double det = (double)a.det[i];
            int fps[3];
            double gl[3], gr[3];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                fps[ax] = s_fps[ax * NS + s];
                gl[ax] = (double)s_gl[ax * NS + s];
                gr[ax] = (double)s_gr[ax * NS + s];
            }
            int ev[kInc];
#pragma unroll
            for (int q = 0; q < kInc; ++q) ev[q] = a.inc[cell * kInc + q];
            double cv[kInc];
#pragma unroll
            for (int q = 0; q < kInc; ++q) {
                const int e = ev[q];
                const int axis = (e & 7) >> 1;
                const int fp = axis == 0 ? fps[0] : (axis == 1 ? fps[1] : fps[2]);
                cv[q] = e < 0 ? 0.0 : (double)a.crec[((int64_t)(e >> 4) * NFP + fp) * 2 + f];
            }
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < kInc; ++q) {
                const int e = ev[q];
                if (e < 0) continue;
                const int role = (e >> 3) & 1;
                const int code = e & 7;
                const int axis = code >> 1;
                const double g = (code & 1) ? (axis == 0 ? gr[0] : (axis == 1 ? gr[1] : gr[2]))
                                            : (axis == 0 ? gl[0] : (axis == 1 ? gl[1] : gl[2]));
                acc += (role == 0 ? -1.0 : 1.0) * g * cv[q];
            }
            double off[5];
            a.so.row(cell, off);
            const double rho = dmax((double)a.U[i * 5] + off[0], 1e-10);
            st[t] = ((double)a.vol[i * 2 + f] + acc / det) / rho;
            sf[t] = fabs((double)(f == 0 ? a.k[i] : a.w[i]));
        }
        wave_sync();
        // outlier zeroing per field (troubled_cell.py:339-408 with n_vars = 1, once for k and once
        // for omega, each field's own mean |value| as the reference magnitude), then non-finite
        // transport values -> 0 (transport.py:509-515)
        outlier_zeroing<NS, 2, double>(st, [sf](int f) {
            double m = 0.0;
            for (int s2 = 0; s2 < NS; ++s2) m += sf[s2 * 2 + f];
            return m / NS;
        }, t);
        wave_sync();
        if (t < 2 * NS && !isfinite(st[t])) st[t] = 0.0;
        wave_sync();
        if (t < NS) {
            const int s = t;
            const int64_t i = cell * NS + s;
            const double dt = (double)a.dt[i];
            const double k = (double)a.k[i], w = (double)a.w[i];
            const double wsafe = dmax(w, 1e-10);
            const double beta = (double)a.src[i * 3 + 2];
            double dk = (double)a.src[i * 3 + 0] / (1.0 + dt * BSTAR * wsafe);
            double dw = (double)a.src[i * 3 + 1] / (1.0 + dt * beta * wsafe);
            dk = finite_or(dk, 0.0) + st[s * 2 + 0];
            dw = finite_or(dw, 0.0) + st[s * 2 + 1];
            dk = finite_or(dk, 0.0);
            dw = finite_or(dw, 0.0);
            if (a.rates) {
                // residual evaluation only: dk/dt and domega/dt of the update, fields untouched
                a.rates[i * 2 + 0] = (real)dk;
                a.rates[i * 2 + 1] = (real)dw;
            } else {
                double kn = k + dt * dk, wn = w + dt * dw;
                kn = isfinite(kn) ? kn : 1e-12;
                wn = isfinite(wn) ? wn : 1e-12;
                a.k[i] = (real)dmax(kn, 1e-12);
                a.w[i] = (real)dmax(wn, 1e-12);
            }
        }
        wave_sync();
    }
}

// nearest wall node (FP64 brute force, wall nodes staged through LDS)
template <typename real>
__global__ void __launch_bounds__(256) k_wall_distance(const double* __restrict__ xyz, int64_t n,
                                                       const double* __restrict__ wall, int n_wall, real* out) {
    __shared__ double s_w[256 * 3];
    for (int64_t base = (int64_t)blockIdx.x * 256; base < n; base += (int64_t)gridDim.x * 256) {
        const int64_t i = base + threadIdx.x;
        double x = 0, y = 0, z = 0;
        if (i < n) {
            x = xyz[i * 3 + 0];
            y = xyz[i * 3 + 1];
            z = xyz[i * 3 + 2];
        }
        double best = 1e300;
        for (int w0 = 0; w0 < n_wall; w0 += 256) {
            __syncthreads();
            const int m = n_wall - w0 < 256 ? n_wall - w0 : 256;
            if (threadIdx.x < m) {
                s_w[threadIdx.x * 3 + 0] = wall[(int64_t)(w0 + threadIdx.x) * 3 + 0];
                s_w[threadIdx.x * 3 + 1] = wall[(int64_t)(w0 + threadIdx.x) * 3 + 1];
                s_w[threadIdx.x * 3 + 2] = wall[(int64_t)(w0 + threadIdx.x) * 3 + 2];
            }
            __syncthreads();
            for (int j = 0; j < m; ++j) {
                const double dx = x - s_w[j * 3 + 0], dy = y - s_w[j * 3 + 1], dz = z - s_w[j * 3 + 2];
                const double d2 = dx * dx + dy * dy + dz * dz;
                best = d2 < best ? d2 : best;
            }
        }
        if (i < n) out[i] = (real)sqrt(best);
    }
}

template <int P, int CT, typename real, int PAIR>
int launch_sst_cell_t(const SstArgs<real>& a, hipStream_t st) {
    constexpr int CPW = SstCellPlan<P, CT, PAIR>::CPW;
    const int64_t units = (a.n_cells + CPW - 1) / CPW;
    hipLaunchKernelGGL((k_sst_cell<P, CT, real, PAIR>),
                       dim3((unsigned)walk_grid(units, kWaves, a.grid_cap > 0 ? a.grid_cap : 4096)), dim3(64 * kWaves),
                       0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int P, int CT, typename real>
int launch_sst_cell(const SstArgs<real>& a, hipStream_t st) {
    if (a.n_cells <= 0) return 0;
    if (a.variant == 0) return launch_sst_cell_t<P, CT, real, 1>(a, st);
    if (a.variant == 1) return launch_sst_cell_t<P, CT, real, 0>(a, st);
    return -1;
}

template <int P, typename real>
int launch_sst_face(const SstArgs<real>& a, hipStream_t st) {
    if (a.n_records <= 0) return 0;
    const int64_t n = a.n_records * (P + 1) * (P + 1);
    const int64_t need = (n + 255) / 256;
    constexpr int KS = (P + 1) * (P + 1), NS = KS * (P + 1);
    const int64_t cap = a.grid_cap_face > 0 ? a.grid_cap_face : 8192;
    const dim3 grid((unsigned)(need < cap ? need : cap));
    if (a.knz == KS)
        hipLaunchKernelGGL((k_sst_face<P, KS, real>), grid, dim3(256), 0, st, a);
    else if (a.knz == NS)
        hipLaunchKernelGGL((k_sst_face<P, NS, real>), grid, dim3(256), 0, st, a);
    else
        return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int P, int CT, typename real>
int launch_sst_update(const SstArgs<real>& a, hipStream_t st) {
    if (a.n_cells <= 0) return 0;
    hipLaunchKernelGGL((k_sst_update<P, CT, real>), dim3((unsigned)walk_grid(a.n_cells, kWaves, a.grid_cap > 0 ? a.grid_cap : 4096)), dim3(64 * kWaves), 0,
                       st, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_wall_distance(const double* sp_xyz, int64_t n_pts, const double* wall_xyz, int n_wall, void* d_out,
                         int dtype, hipStream_t st) {
    const int64_t need = (n_pts + 255) / 256;
    const dim3 grid((unsigned)(need < 16384 ? (need > 0 ? need : 1) : 16384));
    if (dtype == AFR_F32)
        hipLaunchKernelGGL((k_wall_distance<float>), grid, dim3(256), 0, st, sp_xyz, n_pts, wall_xyz, n_wall,
                           (float*)d_out);
    else
        hipLaunchKernelGGL((k_wall_distance<double>), grid, dim3(256), 0, st, sp_xyz, n_pts, wall_xyz, n_wall,
                           (double*)d_out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

#define AFR_SST_INST(P, CT, T)                                                  \
    template int launch_sst_cell<P, CT, T>(const SstArgs<T>&, hipStream_t);     \
    template int launch_sst_update<P, CT, T>(const SstArgs<T>&, hipStream_t);
AFR_SST_INST(1, 0, float)
AFR_SST_INST(2, 0, float)
AFR_SST_INST(1, 1, float)
AFR_SST_INST(2, 1, float)
AFR_SST_INST(1, 0, double)
AFR_SST_INST(2, 0, double)
AFR_SST_INST(1, 1, double)
AFR_SST_INST(2, 1, double)
template int launch_sst_face<1, float>(const SstArgs<float>&, hipStream_t);
template int launch_sst_face<2, float>(const SstArgs<float>&, hipStream_t);
template int launch_sst_face<1, double>(const SstArgs<double>&, hipStream_t);
template int launch_sst_face<2, double>(const SstArgs<double>&, hipStream_t);

}  // namespace afr_sst
