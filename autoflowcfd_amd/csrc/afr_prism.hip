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

typedef __attribute__((address_space(3))) void lds_void_t;
typedef const __attribute__((address_space(1))) void gbl_void_t;

// Asynchronous global -> LDS copy of one dword per lane (gfx950 global_load_lds_dword):
// lane i writes dst[i]; completion is awaited with vm_wait_all().
__device__ __forceinline__ void lds_dma_dword(const void* src_lane, void* dst_wave) {
    __builtin_amdgcn_global_load_lds((gbl_void_t*)src_lane, (lds_void_t*)dst_wave, 4, 0, 0);
}

__device__ __forceinline__ void vm_wait_all() { __builtin_amdgcn_s_waitcnt(0x0F70); /* vmcnt(0) */ }

// Exact median-based outlier zeroing of the (NS x 5) residual block of one
// cell held in s_R ([s][v]) and in rv (lane-strided copy), as in
// troubled_cell.py:339-408: R[s][v] = 0 where |R| > 1e4 max(median_s |R[:,v]|,
// 1e-9 mean_s |U[:,v]|, tiny).  A cheap certificate runs first: with
// M = max_s |R[:,v]| and t = 1.0001e-4 M, if at most k1 = (NS-1)/2 values lie
// below t then median >= t, so no value can exceed the threshold and the
// variable is left as is.  Only when the certificate fails for some variable
// is the median selected exactly (rank counting).
template <int NS, typename real>
__device__ __forceinline__ void outlier_zeroing(const real* s_R, real* s_med, real* rv, const real* umean, int t) {
    constexpr int NR = (NS * 5 + 63) / 64;
    constexpr int k1 = (NS - 1) / 2, k2 = NS / 2;
    constexpr int G = 8, J = (NS + G - 1) / G;  // 8 lanes per variable, J values each
    const int v = t >> 3, g = t & 7;
    real mx = 0;
    if (t < 5 * G) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int s = g + G * j;
            if (s < NS) {
                const real x = fabs(s_R[s * 5 + v]);
                mx = x > mx ? x : mx;
            }
        }
    }
#pragma unroll
    for (int m = 1; m < G; m <<= 1) {
        const real o = __shfl_xor(mx, m);
        mx = o > mx ? o : mx;
    }
    const real tt = mx * real(1.0001e-4);
    int cnt = 0;
    if (t < 5 * G) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int s = g + G * j;
            if (s < NS) cnt += fabs(s_R[s * 5 + v]) < tt;
        }
    }
#pragma unroll
    for (int m = 1; m < G; m <<= 1) cnt += __shfl_xor(cnt, m);
    const bool need = (t < 5 * G) && cnt > k1;
    if (!__any(need)) return;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int o = t + 64 * k;
        if (o < NS * 5) {
            const int vv = o % 5;
            const real x = fabs(rv[k]);
            int lt = 0, le = 0;
            for (int s2 = 0; s2 < NS; ++s2) {
                const real y = fabs(s_R[s2 * 5 + vv]);
                lt += y < x;
                le += y <= x;
            }
            if (lt <= k1 && k1 < le) s_med[vv * 2 + 0] = x;
            if (lt <= k2 && k2 < le) s_med[vv * 2 + 1] = x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int o = t + 64 * k;
        if (o < NS * 5) {
            const int vv = o % 5;
            const real med = (NS % 2) ? s_med[vv * 2] : real(0.5) * (s_med[vv * 2] + s_med[vv * 2 + 1]);
            const real um = real(1e-9) * umean[vv];
            real ref = med > um ? med : um;
            ref = ref > Tiny<real>::v ? ref : Tiny<real>::v;
            if (fabs(rv[k]) > real(1e4) * ref) rv[k] = real(0);
        }
    }
}

template <int P, int KMAX, typename real>
__global__ void __launch_bounds__(64) k_prism_inviscid(PrismArgs<real> a) {
    using K = PK<P>;
    constexpr int N1 = K::N1, NAB = K::NAB, NF1 = K::NF1, NABF = K::NABF, NF = K::NF, NS = K::NS, NFP = K::NFP;
    constexpr int NSV = NS * 5;
    constexpr int NR = (NSV + 63) / 64;
    constexpr int NNB = 10;                             // source cells: 5 slots x 2 sources
    constexpr int WPR = sizeof(real) / 4;               // dwords per real
    constexpr int NDW = NSV * WPR;                      // dwords of one cell's state
    constexpr int NCH = (NDW + 63) / 64;                // 64-dword chunks per cell
    __shared__ double s_ref[20];                        // qref (5), F(qref) (15)
    __shared__ real s_tab[K::TABN];
    __shared__ real s_g[2 * N1];
    __shared__ int s_fp[3 * NS];
    __shared__ int s_co[3 * NS];
    __shared__ int s_nb[NNB];
    __shared__ real s_Q[NSV];
    __shared__ __attribute__((aligned(16))) uint32_t s_N[NNB * NCH * 64];  // source-cell states (raw dwords)
    __shared__ int s_ridx[KMAX * 64];                  // interpolation row of each interface lane
    __shared__ uint32_t s_rw[KMAX * 64 * WPR];
    __shared__ real s_X[NAB * NF1 * 5];
    __shared__ real s_Ft[3 * NF * 5];
    __shared__ real s_YW[(NAB * NF1 + NABF * N1) * 5];
    __shared__ real s_J[5 * NFP * 5];
    __shared__ real s_R[NSV];
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
    if (t < 5) s_ref[t] = a.qref[t];
    if (t < 15) s_ref[5 + t] = a.fref[t];
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

    // interface lane (slot, flux point) bookkeeping
    const bool iface = t < 5 * NFP;
    const int slot = iface ? t / NFP : 0, fp = iface ? t % NFP : 0;
    const uint32_t* dQw = reinterpret_cast<const uint32_t*>(a.dQ);
    const uint32_t* rww = reinterpret_cast<const uint32_t*>(a.rows_w);

    for (int64_t cell = blockIdx.x; cell < a.n_cells; cell += gridDim.x) {
        const int64_t sidx = cell * 5 + slot;
        // 1. delta primitive state of the cell, its source cells
        for (int i = t; i < NSV; i += 64) s_Q[i] = a.dQ[cell * NSV + i];
        if (t < NNB) s_nb[t] = a.slot_nb[cell * NNB + t];
        if (t < 10) s_med[t] = real(0);
        __syncthreads();
        // source-cell states and interpolation rows go to LDS asynchronously (global_load_lds),
        // awaited before the interface stage
        {
#pragma unroll
            for (int j = 0; j < NNB; ++j) {
                const int nbc = s_nb[j];
                const int64_t base = nbc >= 0 ? (int64_t)nbc * NDW : 0;
#pragma unroll
                for (int ch = 0; ch < NCH; ++ch) {
                    const int w = ch * 64 + t;
                    lds_dma_dword(dQw + base + (w < NDW ? w : NDW - 1), s_N + (j * NCH + ch) * 64);
                }
            }
            const int bc0 = iface ? a.slot_bc[sidx] : BC_INTERIOR;
            int64_t row = a.n_rows;  // the all-zero row
            if (iface && bc0 == BC_INTERIOR) {
                const int src = (a.slot_mask[sidx] >> fp) & 1;
                const int r = a.slot_blk[sidx * 2 + src] * NFP + fp;
                if (r >= 0 && r < a.n_rows && s_nb[slot * 2 + src] >= 0) row = r;
            }
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                lds_dma_dword(a.rows_idx + row * a.knz + k, s_ridx + k * 64);
#pragma unroll
                for (int h = 0; h < WPR; ++h)
                    lds_dma_dword(rww + (row * a.knz + k) * WPR + h, s_rw + (k * WPR + h) * 64);
            }
        }
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
            for (int v = 0; v < 5; ++v) q[v] = s_ref[v] + (double)dq[v];
            euler_flux(q, Fd);
            real F[3][5];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int v = 0; v < 5; ++v) F[i][v] = (real)(Fd[i][v] - s_ref[5 + i * 5 + v]);
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
        vm_wait_all();
        __syncthreads();
        // 5a. interface jumps: lane = (slot, flux point)
        if (iface) {
            const int bc = a.slot_bc[sidx];
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
            for (int v = 0; v < 5; ++v) qo[v] = s_ref[v] + (double)dqo[v];
            double tn[3] = {(double)a.slot_tn[sidx * 3 + 0], (double)a.slot_tn[sidx * 3 + 1],
                            (double)a.slot_tn[sidx * 3 + 2]};
            if (bc == BC_INTERIOR) {
                const int src = (a.slot_mask[sidx] >> fp) & 1;
                const real* qs = reinterpret_cast<const real*>(s_N + (slot * 2 + src) * NCH * 64);
                const real* rwv = reinterpret_cast<const real*>(s_rw);
                real dqn[5] = {0, 0, 0, 0, 0};
#pragma unroll
                for (int k = 0; k < KMAX; ++k) {
                    const int sp = s_ridx[k * 64 + t];
                    real w;
                    if constexpr (WPR == 1) {
                        w = rwv[k * 64 + t];
                    } else {
                        w = __hiloint2double((int)s_rw[(k * WPR + 1) * 64 + t], (int)s_rw[(k * WPR) * 64 + t]);
                    }
#pragma unroll
                    for (int v = 0; v < 5; ++v) dqn[v] += w * qs[sp * 5 + v];
                }
#pragma unroll
                for (int v = 0; v < 5; ++v) qn[v] = s_ref[v] + (double)dqn[v];
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
        real rv[NR];
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const int o = t + 64 * k;
            rv[k] = 0;
            if (o < NSV) {
                const int s = o / 5, v = o % 5, ab = s / N1, c = s % N1;
                real acc = 0;
#pragma unroll
                for (int cf = 0; cf < NF1; ++cf) acc += f2c_c[c * NF1 + cf] * s_Y[(ab * NF1 + cf) * 5 + v];
#pragma unroll 4
                for (int abf = 0; abf < NABF; ++abf) acc += f2c_ab[ab * NABF + abf] * s_W[(abf * N1 + c) * 5 + v];
                const real* gc = a.gcl + (cell * NS + s) * 3;
                acc += gc[0] * (real)s_ref[5 + v] + gc[1] * (real)s_ref[10 + v] + gc[2] * (real)s_ref[15 + v];
                rv[k] = acc;
            }
        }
        __syncthreads();
        // 6. interface correction, residual
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const int o = t + 64 * k;
            if (o < NSV) {
                const int s = o / 5, v = o % 5;
                real corr = 0;
#pragma unroll
                for (int sl = 0; sl < 5; ++sl) {
                    const int code = sl < 3 ? sl : (sl == 3 ? 4 : 5);
                    const int axis = code >> 1;
                    const real g = s_g[((code & 1) ? N1 : 0) + s_co[axis * NS + s]];
                    corr += g * s_J[(sl * NFP + s_fp[axis * NS + s]) * 5 + v];
                }
                const real dj = a.det[cell * NS + s];
                rv[k] = -rv[k] / dj - corr / dj;
                s_R[o] = rv[k];
            }
        }
        __syncthreads();
        // 7. outlier suppression
        if (a.suppress) outlier_zeroing<NS, real>(s_R, s_med, rv, a.umean + cell * 5, t);
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const int o = t + 64 * k;
            if (o < NSV) a.R[cell * NSV + o] = rv[k];
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
    const dim3 grid((unsigned)(g > 0 ? g : 1)), block(64);
    // interpolation rows: Kronecker-sparse (prism sources: P2 <= 9, P1 <= 4 non-zeros) or dense (tet sources)
    constexpr int KS = (P + 1) * (P + 1), NS = PK<P>::NS;
    if (a.knz <= KS)
        hipLaunchKernelGGL((k_prism_inviscid<P, KS, real>), grid, block, 0, st, a);
    else if (a.knz <= NS)
        hipLaunchKernelGGL((k_prism_inviscid<P, NS, real>), grid, block, 0, st, a);
    else
        return -1;
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
