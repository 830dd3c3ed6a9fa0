// Prism FR residual kernels, one 64-lane wavefront per cell (gfx950).
//
// Prism operators factor exactly as (a,b)-triangle x c-line Kronecker products
// (the collapsed prism basis is g_ij(a,b) P_k(c), src/autoflowcfd/fr/collapsed_basis.py:197-230),
// so the over-integrated volume term is applied as sum factorisation:
//   X  = c2f_c  . dQ          (along c)        Qf = c2f_ab . X    (along a,b)
//   F~ = adj_fine (F(Q_ref + Qf) - F(Q_ref))   (FP64 point flux, FP32 result)
//   Y  = A_a F~_a + A_b F~_b  (A = f2c_ab D_ab)  W = B_c F~_c      (B = f2c_c D_c)
//   div = f2c_c . Y + f2c_ab . W  (+ constant-state term gcl . F(Q_ref))
// Lanes map to the 64 fine points (P2) for the flux stage, to (slot, flux point)
// for the interface stage and to (solution point, variable) for the rest.
// Reference semantics as in afr_residual.hip (inviscid.py:139-345,
// inviscid_kernel.py:87-284, troubled_cell.py:339-408).
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/afr.h"
#include "afr_physics.hpp"
#include "afr_prism.hpp"

using namespace afr;

namespace afr_prism {

template <int P, typename real>
__global__ void __launch_bounds__(64) k_prism_inviscid(PrismArgs<real> a) {
    using K = PK<P>;
    constexpr int N1 = K::N1, NAB = K::NAB, NF1 = K::NF1, NABF = K::NABF, NF = K::NF, NS = K::NS, NFP = K::NFP;
    __shared__ real s_tab[K::TABN];
    __shared__ real s_g[2 * N1];
    __shared__ int s_fp[3 * NS];
    __shared__ int s_co[3 * NS];
    __shared__ real s_Q[NS * 5];
    __shared__ real s_X[NAB * NF1 * 5];
    __shared__ real s_Ft[3 * NF * 5];
    __shared__ real s_YW[(NAB * NF1 + NABF * N1) * 5];
    __shared__ real s_J[5 * NFP * 5];
    __shared__ real s_R[NS * 5];
    __shared__ real s_med[10];

    const int t = threadIdx.x;
    for (int i = t; i < K::TABN; i += 64) s_tab[i] = a.tab[i];
    for (int i = t; i < 3 * NS; i += 64) {
        s_fp[i] = a.fp_of_sp[i];
        s_co[i] = a.coord_of_sp[i];
    }
    if (t < N1) {
        s_g[t] = a.gL[t];
        s_g[N1 + t] = a.gR[t];
    }
    __syncthreads();
    const real* c2f_ab = s_tab + K::O_C2FAB;
    const real* c2f_c = s_tab + K::O_C2FC;
    const real* A_a = s_tab + K::O_AA;
    const real* A_b = s_tab + K::O_AB;
    const real* B_c = s_tab + K::O_BC;
    const real* f2c_c = s_tab + K::O_F2CC;
    const real* f2c_ab = s_tab + K::O_F2CAB;
    const real* E_ab = s_tab + K::O_EAB;
    const real* e_c = s_tab + K::O_EC;
    real* s_Y = s_YW;
    real* s_W = s_YW + NAB * NF1 * 5;

    for (int64_t cell = blockIdx.x; cell < a.n_cells; cell += gridDim.x) {
        // 1. delta primitive state of the cell
        for (int i = t; i < NS * 5; i += 64) s_Q[i] = a.dQ[cell * NS * 5 + i];
        if (t < 10) s_med[t] = real(0);
        __syncthreads();
        // 2. coarse -> fine along c
        for (int o = t; o < NAB * NF1 * 5; o += 64) {
            const int ab = o / (NF1 * 5), r = o % (NF1 * 5), cf = r / 5, v = r % 5;
            real acc = 0;
#pragma unroll
            for (int c = 0; c < N1; ++c) acc += c2f_c[cf * N1 + c] * s_Q[(ab * N1 + c) * 5 + v];
            s_X[o] = acc;
        }
        __syncthreads();
        // 3. fine points: along (a,b), FP64 point flux about the reference, contravariant flux
        if (t < NF) {
            const int abf = t / NF1, cf = t % NF1;
            real dq[5] = {0, 0, 0, 0, 0};
#pragma unroll
            for (int ab = 0; ab < NAB; ++ab) {
                const real w = c2f_ab[abf * NAB + ab];
#pragma unroll
                for (int v = 0; v < 5; ++v) dq[v] += w * s_X[(ab * NF1 + cf) * 5 + v];
            }
            double q[5], Fd[3][5];
#pragma unroll
            for (int v = 0; v < 5; ++v) q[v] = a.qref[v] + (double)dq[v];
            euler_flux(q, Fd);
            real F[3][5];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int v = 0; v < 5; ++v) F[i][v] = (real)(Fd[i][v] - a.fref[i * 5 + v]);
            const real* ad = a.adjf + (cell * NF + t) * 9;
            real A[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) A[k] = ad[k];
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    s_Ft[((m * NABF + abf) * NF1 + cf) * 5 + v] =
                        A[m * 3 + 0] * F[0][v] + A[m * 3 + 1] * F[1][v] + A[m * 3 + 2] * F[2][v];
        }
        __syncthreads();
        // 4. divergence partial products
        for (int o = t; o < (NAB * NF1 + NABF * N1) * 5; o += 64) {
            real acc = 0;
            if (o < NAB * NF1 * 5) {
                const int ab = o / (NF1 * 5), r = o % (NF1 * 5), cf = r / 5, v = r % 5;
#pragma unroll 4
                for (int abf = 0; abf < NABF; ++abf) {
                    acc += A_a[ab * NABF + abf] * s_Ft[((0 * NABF + abf) * NF1 + cf) * 5 + v];
                    acc += A_b[ab * NABF + abf] * s_Ft[((1 * NABF + abf) * NF1 + cf) * 5 + v];
                }
            } else {
                const int o2 = o - NAB * NF1 * 5;
                const int abf = o2 / (N1 * 5), r = o2 % (N1 * 5), c = r / 5, v = r % 5;
#pragma unroll
                for (int cf = 0; cf < NF1; ++cf) acc += B_c[c * NF1 + cf] * s_Ft[((2 * NABF + abf) * NF1 + cf) * 5 + v];
            }
            s_YW[o] = acc;
        }
        __syncthreads();
        // 5a. interface jumps: lane = (slot, flux point)
        if (t < 5 * NFP) {
            const int slot = t / NFP, fp = t % NFP;
            const int64_t sidx = cell * 5 + slot;
            const int code = slot < 3 ? slot : (slot == 3 ? 4 : 5);
            const double side = (code & 1) ? 1.0 : -1.0;
            real dqo[5] = {0, 0, 0, 0, 0};
            if (slot < 3) {
                const int x = fp / N1, ic = fp % N1;
#pragma unroll
                for (int ab = 0; ab < NAB; ++ab) {
                    const real w = E_ab[(slot * N1 + x) * NAB + ab];
#pragma unroll
                    for (int v = 0; v < 5; ++v) dqo[v] += w * s_Q[(ab * N1 + ic) * 5 + v];
                }
            } else {
#pragma unroll
                for (int c = 0; c < N1; ++c) {
                    const real w = e_c[(slot - 3) * N1 + c];
#pragma unroll
                    for (int v = 0; v < 5; ++v) dqo[v] += w * s_Q[(fp * N1 + c) * 5 + v];
                }
            }
            double qo[5], qn[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) qo[v] = a.qref[v] + (double)dqo[v];
            double tn[3] = {(double)a.slot_tn[sidx * 3 + 0], (double)a.slot_tn[sidx * 3 + 1],
                            (double)a.slot_tn[sidx * 3 + 2]};
            const int bc = a.slot_bc[sidx];
            if (bc == BC_INTERIOR) {
                const int src = (a.slot_mask[sidx] >> fp) & 1;
                const int nbc = a.slot_nb[sidx * 2 + src];
                const int row = a.slot_blk[sidx * 2 + src] * NFP + fp;
                real dqn[5] = {0, 0, 0, 0, 0};
                if (nbc >= 0 && row >= 0 && row < a.n_rows) {
                    const int nnz = a.rows_nnz[row];
                    const real* qsrc = a.dQ + (int64_t)nbc * NS * 5;
                    for (int k = 0; k < nnz; ++k) {
                        const int s = a.rows_idx[row * a.knz + k];
                        const real w = a.rows_w[row * a.knz + k];
#pragma unroll
                        for (int v = 0; v < 5; ++v) dqn[v] += w * qsrc[s * 5 + v];
                    }
                }
#pragma unroll
                for (int v = 0; v < 5; ++v) qn[v] = a.qref[v] + (double)dqn[v];
            } else {
                ghost_state(bc, qo, tn, a.bcp, qn);
            }
            const real* ar = a.slot_adjrow + (sidx * NFP + fp) * 3;
            const double A0 = ar[0], A1 = ar[1], A2 = ar[2];
            double mag = sqrt(A0 * A0 + A1 * A1 + A2 * A2);
            mag = mag > 1e-300 ? mag : 1e-300;
            double d[3];
            if ((a.slot_fb[sidx] >> fp) & 1) {
                d[0] = tn[0]; d[1] = tn[1]; d[2] = tn[2];
            } else {
                const double im = side / mag;
                d[0] = A0 * im; d[1] = A1 * im; d[2] = A2 * im;
            }
            double fc[5], F[3][5];
            ausm_up(qo, qn, d, fc);
            euler_flux(qo, F);
            const double sc = mag * side;
#pragma unroll
            for (int v = 0; v < 5; ++v)
                s_J[(slot * NFP + fp) * 5 + v] = (real)(fc[v] * sc - (A0 * F[0][v] + A1 * F[1][v] + A2 * F[2][v]));
        }
        // 5b. volume divergence per (solution point, variable)
        real rv[(NS * 5 + 63) / 64];
#pragma unroll
        for (int k = 0; k < (NS * 5 + 63) / 64; ++k) {
            const int o = t + 64 * k;
            rv[k] = 0;
            if (o < NS * 5) {
                const int s = o / 5, v = o % 5, ab = s / N1, c = s % N1;
                real acc = 0;
#pragma unroll
                for (int cf = 0; cf < NF1; ++cf) acc += f2c_c[c * NF1 + cf] * s_Y[(ab * NF1 + cf) * 5 + v];
#pragma unroll 4
                for (int abf = 0; abf < NABF; ++abf) acc += f2c_ab[ab * NABF + abf] * s_W[(abf * N1 + c) * 5 + v];
                const real* gc = a.gcl + (cell * NS + s) * 3;
                acc += gc[0] * (real)a.fref[v] + gc[1] * (real)a.fref[5 + v] + gc[2] * (real)a.fref[10 + v];
                rv[k] = acc;
            }
        }
        __syncthreads();
        // 6. interface correction, residual
#pragma unroll
        for (int k = 0; k < (NS * 5 + 63) / 64; ++k) {
            const int o = t + 64 * k;
            if (o < NS * 5) {
                const int s = o / 5, v = o % 5;
                real corr = 0;
#pragma unroll
                for (int slot = 0; slot < 5; ++slot) {
                    const int code = slot < 3 ? slot : (slot == 3 ? 4 : 5);
                    const int axis = code >> 1;
                    const real g = s_g[((code & 1) ? N1 : 0) + s_co[axis * NS + s]];
                    corr += g * s_J[(slot * NFP + s_fp[axis * NS + s]) * 5 + v];
                }
                const real dj = a.det[cell * NS + s];
                rv[k] = -rv[k] / dj - corr / dj;
                s_R[o] = rv[k];
            }
        }
        __syncthreads();
        if (a.suppress) {
            // 7. median of |R| over solution points per variable (parallel rank selection)
            constexpr int k1 = (NS - 1) / 2, k2 = NS / 2;
#pragma unroll
            for (int k = 0; k < (NS * 5 + 63) / 64; ++k) {
                const int o = t + 64 * k;
                if (o < NS * 5) {
                    const int v = o % 5;
                    const real x = rv[k] < real(0) ? -rv[k] : rv[k];
                    int lt = 0, le = 0;
                    for (int s2 = 0; s2 < NS; ++s2) {
                        real y = s_R[s2 * 5 + v];
                        y = y < real(0) ? -y : y;
                        lt += y < x;
                        le += y <= x;
                    }
                    if (lt <= k1 && k1 < le) s_med[v * 2 + 0] = x;
                    if (lt <= k2 && k2 < le) s_med[v * 2 + 1] = x;
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < (NS * 5 + 63) / 64; ++k) {
                const int o = t + 64 * k;
                if (o < NS * 5) {
                    const int v = o % 5;
                    const real med = (NS % 2) ? s_med[v * 2] : real(0.5) * (s_med[v * 2] + s_med[v * 2 + 1]);
                    real ref = med > real(1e-9) * a.umean[cell * 5 + v] ? med : real(1e-9) * a.umean[cell * 5 + v];
                    ref = ref > Tiny<real>::v ? ref : Tiny<real>::v;
                    const real x = rv[k] < real(0) ? -rv[k] : rv[k];
                    if (x > real(1e4) * ref) rv[k] = real(0);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < (NS * 5 + 63) / 64; ++k) {
            const int o = t + 64 * k;
            if (o < NS * 5) a.R[cell * NS * 5 + o] = rv[k];
        }
        __syncthreads();
    }
}

// dQ = primitive(U) - ref and umean[c][v] = mean_s |U[c][s][v]| (outlier field reference)
template <int NS, typename real>
__global__ void __launch_bounds__(256) k_prim_umean(const real* __restrict__ U, real* __restrict__ dQ,
                                                    real* __restrict__ umean, int64_t n_cells, Ref5 ref) {
    constexpr int CPB = 256 / NS;
    __shared__ real s_abs[CPB * NS * 5];
    const int t = threadIdx.x;
    for (int64_t c0 = (int64_t)blockIdx.x * CPB; c0 < n_cells; c0 += (int64_t)gridDim.x * CPB) {
        const int c = t / NS, s = t % NS;
        const bool ok = (t < CPB * NS) && (c0 + c < n_cells);
        if (ok) {
            const int64_t i = (c0 + c) * NS + s;
            double u[5], q[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) u[v] = (double)U[i * 5 + v];
            cons2prim(u, q);
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                dQ[i * 5 + v] = (real)(q[v] - ref.q[v]);
                const real x = U[i * 5 + v];
                s_abs[(c * NS + s) * 5 + v] = x < real(0) ? -x : x;
            }
        }
        __syncthreads();
        if (t < CPB * 5) {
            const int cc = t / 5, v = t % 5;
            if (c0 + cc < n_cells) {
                real acc = 0;
                for (int s2 = 0; s2 < NS; ++s2) acc += s_abs[(cc * NS + s2) * 5 + v];
                umean[(c0 + cc) * 5 + v] = acc / real(NS);
            }
        }
        __syncthreads();
    }
}

template <int P, typename real>
int launch_prism_inviscid(const PrismArgs<real>& a, hipStream_t st) {
    int64_t g = a.n_cells < 65536 ? a.n_cells : 65536;
    hipLaunchKernelGGL((k_prism_inviscid<P, real>), dim3((unsigned)(g > 0 ? g : 1)), dim3(64), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NS, typename real>
int launch_prim_umean(const real* U, real* dQ, real* umean, int64_t n, const double* ref, hipStream_t st) {
    Ref5 r;
    for (int v = 0; v < 5; ++v) r.q[v] = ref ? ref[v] : 0.0;
    constexpr int CPB = 256 / NS;
    int64_t blocks = (n + CPB - 1) / CPB;
    int grid = (int)(blocks < 65536 ? (blocks > 0 ? blocks : 1) : 65536);
    hipLaunchKernelGGL((k_prim_umean<NS, real>), dim3(grid), dim3(256), 0, st, U, dQ, umean, n, r);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template int launch_prism_inviscid<1, float>(const PrismArgs<float>&, hipStream_t);
template int launch_prism_inviscid<2, float>(const PrismArgs<float>&, hipStream_t);
template int launch_prism_inviscid<1, double>(const PrismArgs<double>&, hipStream_t);
template int launch_prism_inviscid<2, double>(const PrismArgs<double>&, hipStream_t);
template int launch_prim_umean<8, float>(const float*, float*, float*, int64_t, const double*, hipStream_t);
template int launch_prim_umean<27, float>(const float*, float*, float*, int64_t, const double*, hipStream_t);
template int launch_prim_umean<64, float>(const float*, float*, float*, int64_t, const double*, hipStream_t);
template int launch_prim_umean<8, double>(const double*, double*, double*, int64_t, const double*, hipStream_t);
template int launch_prim_umean<27, double>(const double*, double*, double*, int64_t, const double*, hipStream_t);
template int launch_prim_umean<64, double>(const double*, double*, double*, int64_t, const double*, hipStream_t);

}  // namespace afr_prism
