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
#include "afr_outlier.hpp"

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

// LDS ordering inside one wavefront: LDS operations of a wave execute in
// program order, so a compiler-level barrier (no s_barrier) separates stages
// whose lanes exchange data through LDS.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The lane index re-materialised per stage: an empty asm that the optimiser must treat
// as redefining it, so per-stage index arithmetic is not hoisted out of the cell
// loop (where it would pin dozens of VGPRs for the whole kernel).
__device__ __forceinline__ int lane_fresh(int t) {
    asm volatile("" : "+v"(t));
    return t;
}

template <typename real> struct V4T;
template <> struct V4T<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct V4T<double> { typedef double type __attribute__((ext_vector_type(4))); };

template <typename real>
__device__ __forceinline__ typename V4T<real>::type ld4(const real* p) {
    return *reinterpret_cast<const typename V4T<real>::type*>(p);
}
template <typename real>
__device__ __forceinline__ void st4(real* p, typename V4T<real>::type v) {
    *reinterpret_cast<typename V4T<real>::type*>(p) = v;
}

constexpr int r4(int n) { return (n + 3) / 4 * 4; }
constexpr int r64(int n) { return (n + 63) / 64 * 64; }

// STAGE (the kernel's tile variant) 1: the five slot source cells' dQ are copied into LDS with
// the own cell's; 0: only the own cell is staged and the interface lanes read the sources from
// global memory.  0 frees 5 x 540 B of LDS per wave at P2 FP32 (29 instead of 39 KB per
// workgroup, 5 instead of 4 waves per SIMD): measured 3.08 -> 2.76 ms on the 82^3 P2 mesh; the
// autotuner picks per mesh.
// Padded LDS layout of the prism kernel (units: reals).
template <int P, typename real, int STAGE> struct PL {
    using K = PK<P>;
    static constexpr int N1 = K::N1, NAB = K::NAB, NF1 = K::NF1, NABF = K::NABF, NF = K::NF, NS = K::NS;
    static constexpr int NFP = K::NFP, NSV = NS * 5;
    static constexpr int NABF4 = r4(NABF), NABP4 = r4(NAB), NG = (NAB + 2) / 3, NX = r4(NAB * 5);
    static constexpr int WPR = sizeof(real) / 4;
    // block-shared operator tables
    static constexpr int T_C2FAB = 0;                          // [NABF][NABP4]   c2f_ab rows
    static constexpr int T_A = T_C2FAB + NABF * NABP4;         // [2][NG][NABF4][4] A_m[3g+r][abf]
    static constexpr int T_F2CAB = T_A + 2 * NG * NABF4 * 4;   // [NG][NABF4][4]  f2c_ab[3g+r][abf]
    static constexpr int T_BC = T_F2CAB + NG * NABF4 * 4;      // [N1][4]         B_c[c][cf]
    static constexpr int T_F2CC = T_BC + N1 * 4;               // [N1][4]         f2c_c[c][cf]
    static constexpr int T_C2FC = T_F2CC + N1 * 4;             // [NF1][N1]
    static constexpr int T_EAB = T_C2FC + NF1 * N1;            // [3][N1][NAB]
    static constexpr int T_EC = T_EAB + 3 * N1 * NAB;          // [2][N1]
    static constexpr int T_G = T_EC + 2 * N1;                  // [2][N1] gL, gR
    static constexpr int T_N = r4(T_G + 2 * N1);
    // per-wave buffers
    static constexpr int QN_DW = r64((STAGE ? 6 : 1) * NSV * WPR);  // own cell (+ 5 slot sources), raw dwords
    static constexpr int W_QN = 0;
    static constexpr int W_XY = W_QN + QN_DW / WPR;            // X [NF1][NX] | Y [NAB][5][4]
    static constexpr int XY_N = r4(NF1 * NX > NAB * 20 ? NF1 * NX : NAB * 20);
    static constexpr int W_FT = W_XY + XY_N;                   // Ft [3][NF1][5][NABF4] | J, R
    static constexpr int FT_N = 3 * NF1 * 5 * NABF4;
    static constexpr int W_J = W_FT;                           // J [5][NFP][5]
    static constexpr int W_R = W_FT + r4(5 * NFP * 5);         // R [NS][5]
    static constexpr int W_W = W_FT + FT_N;                    // W [N1][5][NABF4]
    static constexpr int W_N = W_W + N1 * 5 * NABF4;
    static_assert(r4(5 * NFP * 5) + NSV <= FT_N, "J/R alias does not fit");
};

constexpr int kWaves = 4;  // wavefronts per workgroup


// Copies the states of up to 6 cells (dword offsets base[0..5] into src, each
// CELL_DW dwords) back to back into dst[0 .. NCH*64) with one global_load_lds
// per 64-dword chunk; the chunk -> cell split is resolved at compile time.
template <int CH, int NCH, int CELL_DW>
struct DmaCells {
    __device__ __forceinline__ static void run(const uint32_t* src, const int64_t (&base)[6], uint32_t* dst, int t) {
        constexpr int w0 = CH * 64;
        int64_t b;
        int off;
        if constexpr (CELL_DW >= 64) {
            constexpr int j0 = w0 / CELL_DW, bl = (j0 + 1) * CELL_DW - w0;
            if constexpr (bl >= 64) {
                b = base[j0];
                off = w0 - j0 * CELL_DW + t;
            } else {
                const bool second = t >= bl;
                if constexpr (j0 + 1 < 6) {
                    b = second ? base[j0 + 1] : base[j0];
                    off = second ? t - bl : w0 - j0 * CELL_DW + t;
                } else {
                    b = base[j0];
                    off = second ? 0 : w0 - j0 * CELL_DW + t;
                }
            }
        } else {
            const int w = w0 + t, j = w / CELL_DW;
            off = w - j * CELL_DW;
            b = base[0];
#pragma unroll
            for (int q = 1; q < 6; ++q)
                if (j == q) b = base[q];
            if (j >= 6) off = 0;
        }
        lds_dma_dword(src + b + off, dst + w0);
        if constexpr (CH + 1 < NCH) DmaCells<CH + 1, NCH, CELL_DW>::run(src, base, dst, t);
    }
};

template <int NCH, int CELL_DW>
__device__ __forceinline__ void dma_cells(const uint32_t* src, const int64_t (&base)[6], uint32_t* dst, int t) {
    DmaCells<0, NCH, CELL_DW>::run(src, base, dst, t);
}

template <int P, int KMAX, typename real, int STAGE>
__global__ void __launch_bounds__(64 * kWaves) AFR_WPE_SEL(AFR_WPE_PRISM) k_prism_inviscid(PrismArgs<real> a) {
    using L = PL<P, real, STAGE>;
    using v4 = typename V4T<real>::type;
    constexpr int N1 = L::N1, NAB = L::NAB, NF1 = L::NF1, NABF = L::NABF, NF = L::NF, NS = L::NS;
    constexpr int NFP = L::NFP, NSV = L::NSV, NABF4 = L::NABF4, NABP4 = L::NABP4, NG = L::NG, NX = L::NX;
    constexpr int WPR = L::WPR;
    constexpr int CELL_DW = NSV * WPR;
    __shared__ __attribute__((aligned(16))) real s_tab[L::T_N];
    __shared__ __attribute__((aligned(16))) real s_wave[kWaves][L::W_N];
    __shared__ double s_ref[20];  // qref (5), F(qref) (15)
    __shared__ real s_c[6];       // qref (5) and its total enthalpy H0 in the storage precision

    const int tid = threadIdx.x, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    int t = tid & 63;
    // operator tables in padded layouts (sources: packed factors, PK offsets)
    {
        using K = PK<P>;
        const real* tb = a.tab;
        for (int i = tid; i < L::T_N; i += 64 * kWaves) s_tab[i] = real(0);
        __syncthreads();
        for (int i = tid; i < NABF * NAB; i += 64 * kWaves) {
            const int abf = i / NAB, ab = i % NAB;
            s_tab[L::T_C2FAB + abf * NABP4 + ab] = tb[K::O_C2FAB + i];
        }
        for (int i = tid; i < NAB * NABF; i += 64 * kWaves) {
            const int ab = i / NABF, abf = i % NABF, g = ab / 3, r = ab % 3;
            s_tab[L::T_A + ((0 * NG + g) * NABF4 + abf) * 4 + r] = tb[K::O_AA + i];
            s_tab[L::T_A + ((1 * NG + g) * NABF4 + abf) * 4 + r] = tb[K::O_AB + i];
            s_tab[L::T_F2CAB + (g * NABF4 + abf) * 4 + r] = tb[K::O_F2CAB + i];
        }
        for (int i = tid; i < N1 * NF1; i += 64 * kWaves) {
            const int c = i / NF1, cf = i % NF1;
            s_tab[L::T_BC + c * 4 + cf] = tb[K::O_BC + i];
            s_tab[L::T_F2CC + c * 4 + cf] = tb[K::O_F2CC + i];
            s_tab[L::T_C2FC + i] = tb[K::O_C2FC + i];  // [NF1][N1] has N1*NF1 entries too
        }
        for (int i = tid; i < 3 * N1 * NAB; i += 64 * kWaves) s_tab[L::T_EAB + i] = tb[K::O_EAB + i];
        for (int i = tid; i < 2 * N1; i += 64 * kWaves) s_tab[L::T_EC + i] = tb[K::O_EC + i];
        if (tid < N1) {
            s_tab[L::T_G + tid] = a.gL[tid];
            s_tab[L::T_G + N1 + tid] = a.gR[tid];
        }
        if (tid < 5) s_ref[tid] = a.qref[tid];
        if (tid < 15) s_ref[5 + tid] = a.fref[tid];
        if (tid < 5) s_c[tid] = (real)a.qref[tid];
        if (tid == 0) {
            const double* q = a.qref;
            s_c[5] = (real)(kGamma / (kGamma - 1.0) * q[4] / q[0] + 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]));
        }
        __syncthreads();
    }
    real* const W = s_wave[wid];
    uint32_t* const sQNw = reinterpret_cast<uint32_t*>(W + L::W_QN);
    const real* const sQ = W + L::W_QN;  // own cell [NS][5]; source j at sQ + (1 + j) * NSV
    real* const sX = W + L::W_XY;
    real* const sY = W + L::W_XY;
    real* const sFt = W + L::W_FT;
    real* const sJ = W + L::W_J;
    real* const sR = W + L::W_R;
    real* const sW = W + L::W_W;
    const uint32_t* dQw = reinterpret_cast<const uint32_t*>(a.dQ);

    // XCD-aware schedule: workgroups are dispatched round-robin over the 8 XCDs; each
    // XCD walks one contiguous range of the locality-ordered cell list so that the
    // source cells of neighbouring cells are shared in its L2.
    const int64_t n = a.n_cells;
    const int nxcd = gridDim.x >= 8 ? 8 : 1;
    const int xcd = blockIdx.x % nxcd, lb = blockIdx.x / nxcd;
    const int64_t waves_x = (int64_t)(gridDim.x / nxcd) * kWaves;
    const int64_t lo = n * xcd / nxcd, hi = n * (xcd + 1) / nxcd;

    for (int64_t it = lo + (int64_t)lb * kWaves + wid; it < hi; it += waves_x) {
        const int64_t cell = __builtin_amdgcn_readfirstlane(a.schedule[it]);
        t = lane_fresh(t);
        // ---- 1. own state and slot source states -> LDS (asynchronous copies)
        int nbv = -1;
        if (t < 10) nbv = a.slot_nb[cell * 10 + t];
        int64_t base[6];
        base[0] = cell * CELL_DW;
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const int nb = __builtin_amdgcn_readlane(nbv, 2 * j);
            base[1 + j] = nb >= 0 ? (int64_t)nb * CELL_DW : cell * CELL_DW;
        }
        // the interface lanes' slot tables, loaded with the own cell (one memory round trip)
        int m_bc = BC_INTERIOR, m_mask = 0, m_blk0 = -1, m_blk1 = -1, m_nb0 = -1, m_nb1 = -1, m_fb = 0;
        real m_tn[3] = {0, 0, 0}, m_ar[3] = {0, 0, 0};
        if (t < 5 * NFP) {
            const int64_t sidx = cell * 5 + t / NFP;
            m_bc = a.slot_bc[sidx];
            m_mask = a.slot_mask[sidx];
            m_blk0 = a.slot_blk[sidx * 2];
            m_blk1 = a.slot_blk[sidx * 2 + 1];
            m_nb0 = a.slot_nb[sidx * 2];
            m_nb1 = a.slot_nb[sidx * 2 + 1];
            m_fb = a.slot_fb[sidx];
            const real* ar = a.slot_adjrow + (sidx * NFP + t % NFP) * 3;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                m_tn[q] = a.slot_tn[sidx * 3 + q];
                m_ar[q] = ar[q];
            }
        }
        dma_cells<L::QN_DW / 64, CELL_DW>(dQw, base, sQNw, t);
        vm_wait_all();
        wave_sync();
        // ---- 2. coarse -> fine along c: X[cf][ab*5+v]
        t = lane_fresh(t);
        if (t < NAB * 5) {
            const int ab = t / 5, v = t % 5;
            real q[N1];
#pragma unroll
            for (int c = 0; c < N1; ++c) q[c] = sQ[(ab * N1 + c) * 5 + v];
#pragma unroll
            for (int cf = 0; cf < NF1; ++cf) {
                real acc = 0;
#pragma unroll
                for (int c = 0; c < N1; ++c) acc += s_tab[L::T_C2FC + cf * N1 + c] * q[c];
                sX[cf * NX + t] = acc;
            }
        } else if (t < NX) {
#pragma unroll
            for (int cf = 0; cf < NF1; ++cf) sX[cf * NX + t] = real(0);
        }
        wave_sync();
        // ---- 3. fine points (abf, cf): along (a,b), FP64 point flux, contravariant flux
        t = lane_fresh(t);
        if (t < NF) {
            const int abf = t / NF1, cf = t % NF1;
            real dq[5] = {0, 0, 0, 0, 0};
#pragma unroll
            for (int qa = 0; qa < NABP4 / 4; ++qa) {
                const v4 w = ld4(s_tab + L::T_C2FAB + abf * NABP4 + 4 * qa);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int ab = 4 * qa + e;
                    if (ab < NAB) {
#pragma unroll
                        for (int v = 0; v < 5; ++v) dq[v] += w[e] * sX[cf * NX + ab * 5 + v];
                    }
                }
            }
            // F~_m = adj_m . (F(q_ref + dq) - F(q_ref)) by the exact algebraic expansion of the
            // Euler flux about q_ref: every term carries a perturbation factor, so the
            // difference is formed without cancellation in the storage precision
            const real* ad = a.adjf + (cell * NF + t) * 9;
            const real r0 = s_c[0], u0 = s_c[1], v0 = s_c[2], w0 = s_c[3], p0 = s_c[4], H0 = s_c[5];
            const real dr = dq[0], du = dq[1], dv = dq[2], dw = dq[3], dp = dq[4];
            // dH = g/(g-1) d(p/rho) + d(|u|^2/2)
            const real dpr = (dp * r0 - p0 * dr) * recip(r0 * (r0 + dr));
            const real dH = real(kGamma / (kGamma - 1.0)) * dpr + (u0 * du + v0 * dv + w0 * dw) +
                            real(0.5) * (du * du + dv * dv + dw * dw);
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const real a0 = ad[m * 3 + 0], a1 = ad[m * 3 + 1], a2 = ad[m * 3 + 2];
                const real un0 = u0 * a0 + v0 * a1 + w0 * a2;
                const real dun = du * a0 + dv * a1 + dw * a2;
                const real m0 = r0 * un0;
                const real dm = r0 * dun + dr * (un0 + dun);
                real* o = sFt + (m * NF1 + cf) * 5 * NABF4 + abf;
                o[0 * NABF4] = dm;
                o[1 * NABF4] = m0 * du + dm * (u0 + du) + dp * a0;
                o[2 * NABF4] = m0 * dv + dm * (v0 + dv) + dp * a1;
                o[3 * NABF4] = m0 * dw + dm * (w0 + dw) + dp * a2;
                o[4 * NABF4] = m0 * dH + dm * (H0 + dH);
            }
        } else if (t < NF + (NABF4 - NABF) * NF1) {
            const int abf = NABF + (t - NF) / NF1, cf = (t - NF) % NF1;
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
                for (int v = 0; v < 5; ++v) sFt[((m * NF1 + cf) * 5 + v) * NABF4 + abf] = real(0);
        }
        wave_sync();
        // ---- 4a. Y[ab][v][cf] = sum_abf A_a[ab][abf] Ft_a + A_b[ab][abf] Ft_b  (lanes: row group, cf, v)
        t = lane_fresh(t);
        if (t < NG * NF1 * 5) {
            const int g = t / (NF1 * 5), r = t % (NF1 * 5), cf = r / 5, v = r % 5;
            real acc[3] = {0, 0, 0};
#pragma unroll
            for (int m = 0; m < 2; ++m) {
#pragma unroll
                for (int qa = 0; qa < NABF4 / 4; ++qa) {
                    const v4 f = ld4(sFt + ((m * NF1 + cf) * 5 + v) * NABF4 + 4 * qa);
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const v4 c = ld4(s_tab + L::T_A + ((m * NG + g) * NABF4 + 4 * qa + e) * 4);
                        acc[0] += c[0] * f[e];
                        acc[1] += c[1] * f[e];
                        acc[2] += c[2] * f[e];
                    }
                }
            }
#pragma unroll
            for (int rr = 0; rr < 3; ++rr) {
                const int ab = 3 * g + rr;
                if (ab < NAB) {
                    sY[(ab * 5 + v) * 4 + cf] = acc[rr];
                    if (cf == NF1 - 1)
                        for (int cp = NF1; cp < 4; ++cp) sY[(ab * 5 + v) * 4 + cp] = real(0);
                }
            }
        }
        // ---- 4b. W[c][v][abf] = sum_cf B_c[c][cf] Ft_c[cf][v][abf]  (lanes: v, abf quad, c)
        t = lane_fresh(t);
        constexpr int QB = NABF4 / 4;
        if (t < 5 * QB * N1) {
            const int v = t / (QB * N1), r = t % (QB * N1), qa = r / N1, c = r % N1;
            const v4 bc = ld4(s_tab + L::T_BC + c * 4);
            v4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int cf = 0; cf < NF1; ++cf) {
                const v4 f = ld4(sFt + ((2 * NF1 + cf) * 5 + v) * NABF4 + 4 * qa);
                acc += bc[cf] * f;
            }
            st4(sW + (c * 5 + v) * NABF4 + 4 * qa, acc);
        }
        wave_sync();
        // ---- 5a. interface jumps: lane = (slot, flux point)
        t = lane_fresh(t);
        if (t < 5 * NFP) {
            const int slot = t / NFP, fp = t % NFP;
            const int64_t sidx = cell * 5 + slot;
            const int bc = m_bc;
            const int code = slot < 3 ? slot : (slot == 3 ? 4 : 5);
            const double side = (code & 1) ? 1.0 : -1.0;
            real dqo[5] = {0, 0, 0, 0, 0};
            if (slot < 3) {
                const int x = fp / N1, ic = fp % N1;
#pragma unroll
                for (int ab = 0; ab < NAB; ++ab) {
                    const real w = s_tab[L::T_EAB + (slot * N1 + x) * NAB + ab];
#pragma unroll
                    for (int v = 0; v < 5; ++v) dqo[v] += w * sQ[(ab * N1 + ic) * 5 + v];
                }
            } else {
#pragma unroll
                for (int c = 0; c < N1; ++c) {
                    const real w = s_tab[L::T_EC + (slot - 3) * N1 + c];
#pragma unroll
                    for (int v = 0; v < 5; ++v) dqo[v] += w * sQ[(fp * N1 + c) * 5 + v];
                }
            }
            double qo[5], qn[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) qo[v] = s_ref[v] + (double)dqo[v];
            double tn[3] = {(double)m_tn[0], (double)m_tn[1], (double)m_tn[2]};
            if (bc == BC_INTERIOR) {
                const int src = (m_mask >> fp) & 1;
                int row = (src ? m_blk1 : m_blk0) * NFP + fp;
                const int nbc = src ? m_nb1 : m_nb0;
                if (row < 0 || row >= a.n_rows || nbc < 0) row = a.n_rows;  // all-zero row
                int sp[KMAX];
                real rw[KMAX];
#pragma unroll
                for (int k = 0; k < KMAX; ++k) {  // rows are KMAX wide (zero-padded)
                    sp[k] = a.rows_idx[(int64_t)row * KMAX + k];
                    rw[k] = a.rows_w[(int64_t)row * KMAX + k];
                }
                real dqn[5] = {0, 0, 0, 0, 0};
                if (STAGE && src == 0) {  // primary source: staged in LDS
                    const real* qs = sQ + (1 + slot) * NSV;
#pragma unroll
                    for (int k = 0; k < KMAX; ++k)
#pragma unroll
                        for (int v = 0; v < 5; ++v) dqn[v] += rw[k] * qs[sp[k] * 5 + v];
                } else {  // second source of a split quad: global memory
                    const real* qs = a.dQ + (int64_t)(nbc >= 0 ? nbc : 0) * NSV;
#pragma unroll
                    for (int k = 0; k < KMAX; ++k)
#pragma unroll
                        for (int v = 0; v < 5; ++v) dqn[v] += rw[k] * qs[sp[k] * 5 + v];
                }
#pragma unroll
                for (int v = 0; v < 5; ++v) qn[v] = s_ref[v] + (double)dqn[v];
            } else {
                ghost_state(bc, qo, tn, a.bcp + (a.cell_case ? 40 * a.cell_case[cell] : 0), qn);
            }
            const double A0 = m_ar[0], A1 = m_ar[1], A2 = m_ar[2];
            double mag = sqrt_r(A0 * A0 + A1 * A1 + A2 * A2);
            mag = mag > 1e-300 ? mag : 1e-300;
            double d[3];
            if ((m_fb >> fp) & 1) {
                d[0] = tn[0]; d[1] = tn[1]; d[2] = tn[2];
            } else {
                const double im = side * recip(mag);
                d[0] = A0 * im; d[1] = A1 * im; d[2] = A2 * im;
            }
            double fc[5], FA[5];
            ausm_up(qo, qn, d, fc);
            const double Av[3] = {A0, A1, A2};
            euler_flux_dir(qo, Av, FA);
            const double sc = mag * side;
#pragma unroll
            for (int v = 0; v < 5; ++v) sJ[(slot * NFP + fp) * 5 + v] = (real)(fc[v] * sc - FA[v]);
        }
        wave_sync();
        // ---- 5b/6. volume divergence + constant-state term + interface correction, / det
        t = lane_fresh(t);
        //      lanes: (c, v, row group g), rows ab = 3g..3g+2
        if (t < N1 * 5 * NG) {
            const int c = t / (5 * NG), r = t % (5 * NG), v = r / NG, g = r % NG;
            real acc[3] = {0, 0, 0};
            const v4 fc4 = ld4(s_tab + L::T_F2CC + c * 4);
#pragma unroll
            for (int rr = 0; rr < 3; ++rr) {
                const int ab = 3 * g + rr;
                if (ab < NAB) {
                    const v4 y = ld4(sY + (ab * 5 + v) * 4);
                    acc[rr] = fc4[0] * y[0] + fc4[1] * y[1] + fc4[2] * y[2] + fc4[3] * y[3];
                }
            }
#pragma unroll
            for (int qa = 0; qa < NABF4 / 4; ++qa) {
                const v4 w = ld4(sW + (c * 5 + v) * NABF4 + 4 * qa);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const v4 cf = ld4(s_tab + L::T_F2CAB + (g * NABF4 + 4 * qa + e) * 4);
                    acc[0] += cf[0] * w[e];
                    acc[1] += cf[1] * w[e];
                    acc[2] += cf[2] * w[e];
                }
            }
            const real fr0 = (real)s_ref[5 + v], fr1 = (real)s_ref[10 + v], fr2 = (real)s_ref[15 + v];
#pragma unroll
            for (int rr = 0; rr < 3; ++rr) {
                const int ab = 3 * g + rr;
                if (ab < NAB) {
                    const int i = ab / N1, j = ab % N1, s = ab * N1 + c;
                    const real* gc = a.gcl + (cell * NS + s) * 3;
                    const real vol = acc[rr] + gc[0] * fr0 + gc[1] * fr1 + gc[2] * fr2;
                    const real* gL = s_tab + L::T_G;
                    const real* gR = s_tab + L::T_G + N1;
                    const real corr = gL[i] * sJ[(0 * NFP + j * N1 + c) * 5 + v] +
                                      gR[i] * sJ[(1 * NFP + j * N1 + c) * 5 + v] +
                                      gL[j] * sJ[(2 * NFP + i * N1 + c) * 5 + v] +
                                      gL[c] * sJ[(3 * NFP + i * N1 + j) * 5 + v] +
                                      gR[c] * sJ[(4 * NFP + i * N1 + j) * 5 + v];
                    sR[s * 5 + v] = -(vol + corr) * recip(a.det[cell * NS + s]);
                }
            }
        }
        wave_sync();
        // ---- 7. outlier suppression, coalesced write
        t = lane_fresh(t);
        if (a.suppress) {
            const real* um = a.umean + cell * 5;
            outlier_zeroing<NS, 5, real>(sR, [um](int v) { return um[v]; }, t);
        }
        wave_sync();
        for (int o = t; o < NSV; o += 64) a.R[cell * NSV + o] = sR[o];
        wave_sync();
    }
}

// dQ = primitive(U) - ref and umean[c][v] = mean_s |U[c][s][v]| (outlier field reference)
template <int NS, typename real>
__global__ void __launch_bounds__(256) k_prim_umean(const real* __restrict__ U, real* __restrict__ dQ,
                                                    real* __restrict__ umean, int64_t n_cells, Ref5 ref,
                                                    StateOff so) {
    constexpr int CPB = 256 / NS;
    __shared__ real s_abs[CPB * NS * 5];
    const int t = threadIdx.x;
    for (int64_t c0 = (int64_t)blockIdx.x * CPB; c0 < n_cells; c0 += (int64_t)gridDim.x * CPB) {
        const int c = t / NS, s = t % NS;
        const bool ok = (t < CPB * NS) && (c0 + c < n_cells);
        if (ok) {
            const int64_t i = (c0 + c) * NS + s;
            double u[5], q[5];
            load_state(U, i, so, c0 + c, u);
            cons2prim(u, q);
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                dQ[i * 5 + v] = (real)(q[v] - ref.q[v]);
                const real x = (real)u[v];
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
    // persistent grid: a multiple of the 8 XCDs, up to 4 workgroups per CU of 256 CUs
    const int64_t need = (a.n_cells + kWaves - 1) / kWaves;
    const int64_t cap = a.grid_cap > 0 ? a.grid_cap : 1024;
    int64_t g = need < cap ? need : cap;
    g = g >= 8 ? g / 8 * 8 : g;
    const dim3 grid((unsigned)(g > 0 ? g : 1)), block(64 * kWaves);
    // interpolation rows: Kronecker-sparse (prism sources: P2 <= 9, P1 <= 4 non-zeros) or dense (tet sources)
    constexpr int KS = (P + 1) * (P + 1), NS = PK<P>::NS;
    if (a.variant < 0 || a.variant > 1) return -1;
    if (a.knz == KS) {
        if (a.variant) hipLaunchKernelGGL((k_prism_inviscid<P, KS, real, 1>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_prism_inviscid<P, KS, real, 0>), grid, block, 0, st, a);
    } else if (a.knz == NS) {
        if (a.variant) hipLaunchKernelGGL((k_prism_inviscid<P, NS, real, 1>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_prism_inviscid<P, NS, real, 0>), grid, block, 0, st, a);
    } else {
        return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NS, typename real>
int launch_prim_umean(const real* U, real* dQ, real* umean, int64_t n, const double* ref, StateOff so,
                      hipStream_t st) {
    Ref5 r;
    for (int v = 0; v < 5; ++v) r.q[v] = ref ? ref[v] : 0.0;
    constexpr int CPB = 256 / NS;
    int64_t blocks = (n + CPB - 1) / CPB;
    int grid = (int)(blocks < 65536 ? (blocks > 0 ? blocks : 1) : 65536);
    hipLaunchKernelGGL((k_prim_umean<NS, real>), dim3(grid), dim3(256), 0, st, U, dQ, umean, n, r, so);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template int launch_prism_inviscid<1, float>(const PrismArgs<float>&, hipStream_t);
template int launch_prism_inviscid<2, float>(const PrismArgs<float>&, hipStream_t);
template int launch_prism_inviscid<1, double>(const PrismArgs<double>&, hipStream_t);
template int launch_prism_inviscid<2, double>(const PrismArgs<double>&, hipStream_t);
template int launch_prim_umean<8, float>(const float*, float*, float*, int64_t, const double*, StateOff,
                                                 hipStream_t);
template int launch_prim_umean<27, float>(const float*, float*, float*, int64_t, const double*, StateOff,
                                                 hipStream_t);
template int launch_prim_umean<64, float>(const float*, float*, float*, int64_t, const double*, StateOff,
                                                 hipStream_t);
template int launch_prim_umean<8, double>(const double*, double*, double*, int64_t, const double*, StateOff,
                                                 hipStream_t);
template int launch_prim_umean<27, double>(const double*, double*, double*, int64_t, const double*, StateOff,
                                                 hipStream_t);
template int launch_prim_umean<64, double>(const double*, double*, double*, int64_t, const double*, StateOff,
                                                 hipStream_t);

}  // namespace afr_prism
