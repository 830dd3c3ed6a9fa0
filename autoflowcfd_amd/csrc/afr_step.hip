// One steady pseudo-step on the device: local time step, SSP-RK3 stages with
// positivity projection and modal filter, FP64 residual norm.
//
//   k_dt_local     src/autoflowcfd/core/fr_solver/cfl.py:12-147 (primitives with the
//                  floors of src/autoflowcfd/core/fr_solver/state.py:76-97)
//   k_rk_stage     src/autoflowcfd/core/time_integration/base.py:250-335 (Shu-Osher
//                  stage), positivity.py:13-45, src/autoflowcfd/core/fr_solver/filter.py:29-43
//   k_sq_partial / k_sum_partials
//                  src/autoflowcfd/core/fr_solver/state.py:99-107 (RMS, FP64 accumulation,
//                  fixed summation order: deterministic)
#include <hip/hip_runtime.h>

#include "../../include/afr.h"
#include "afr_physics.hpp"
#include "afr_step.hpp"

using namespace afr;

namespace afr_step {

constexpr int kThreads = 256;

// local time step of one solution point i of cell c (h = cbrt |vol[c]|)
template <typename real>
__device__ __forceinline__ real dt_point(const real* __restrict__ U, const StateOff& so, const real* __restrict__ mu_t,
                                         const real* __restrict__ det, const real* __restrict__ mfs, int64_t i,
                                         int64_t c, double h, double mu, int order, int nut_mode) {
    const double cfl = 0.1, fa = 1.0 / (2 * order + 1), fv = fa * fa;
    double u5[5];
    load_state(U, i, so, c, u5);
    const double rho = u5[0] > 1e-10 ? u5[0] : 1e-10;
    const double ir = 1.0 / rho;
    const double u = u5[1] * ir, v = u5[2] * ir, w = u5[3] * ir;
    const double ke = 0.5 * (u * u + v * v + w * w);
    double p = (u5[4] - rho * ke) * 0.4;
    p = p > 1.0 ? p : 1.0;
    double a2 = 1.4 * p / rho;
    a2 = a2 > 1e-10 ? a2 : 1e-10;
    double ws = sqrt(u * u + v * v + w * w) + sqrt(a2);
    ws = ws > 1e-10 ? ws : 1e-10;
    const double dta = cfl * fa * h / ws;
    // nut_mode: the eddy-viscosity array holds nu_t and mu_t = rho nu_t
    double me = mu + (mu_t ? (nut_mode ? rho * (double)mu_t[i] : (double)mu_t[i]) : 0.0);
    me = me > 1e-30 ? me : 1e-30;
    const double dtv = 0.25 * cfl * fv * rho * h * h / me;
    double den = (double)mfs[i] * ws;
    den = den > 1e-300 ? den : 1e-300;
    const double dtg = cfl * fabs((double)det[i]) / den;
    double d = dta < dtv ? dta : dtv;
    d = d < dtg ? d : dtg;
    return (real)d;
}

// any points per cell: one thread per point, grid-stride
template <typename real>
__global__ void __launch_bounds__(kThreads) k_dt_local(const real* __restrict__ U, StateOff so,
                                                       const real* __restrict__ mu_t,
                                                       const real* __restrict__ det, const real* __restrict__ mfs,
                                                       const real* __restrict__ vol, real* __restrict__ dt,
                                                       int64_t n_cells, int ns, double mu, int order, int nut_mode) {
    const int64_t n = n_cells * ns;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = i / ns;
        dt[i] = dt_point(U, so, mu_t, det, mfs, i, c, cbrt(fabs((double)vol[c])), mu, order, nut_mode);
    }
}

// NS points per cell: a workgroup takes CPB = 256 / NS consecutive cells at a time, one thread
// per point (a 32-bit division by the constant NS instead of a 64-bit one; no barrier, so the
// loads of successive passes overlap; the same values as k_dt_local)
template <int NS, typename real>
__global__ void __launch_bounds__(kThreads) k_dt_cells(const real* __restrict__ U, StateOff so,
                                                       const real* __restrict__ mu_t,
                                                       const real* __restrict__ det, const real* __restrict__ mfs,
                                                       const real* __restrict__ vol, real* __restrict__ dt,
                                                       int64_t n_cells, double mu, int order, int nut_mode) {
    constexpr int CPB = kThreads / NS;
    const int t = threadIdx.x;
    if (t >= CPB * NS) return;
    const int lc = t / NS;
    for (int64_t cb = (int64_t)blockIdx.x * CPB; cb < n_cells; cb += (int64_t)gridDim.x * CPB) {
        const int64_t c = cb + lc;
        if (c < n_cells) {
            const int64_t i = cb * NS + t;
            dt[i] = dt_point(U, so, mu_t, det, mfs, i, c, cbrt(fabs((double)vol[c])), mu, order, nut_mode);
        }
    }
}

// U_out = c0 U0 + c1 Ua + c2 dt R, then positivity (rho >= 1e-6, |u| <= 1e4,
// p >= p_floor) per solution point and the modal filter per cell.  A workgroup
// takes CPB = 256 / NS consecutive cells at a time: one thread per solution point
// forms the stage value and its projection into LDS, then one thread per output
// value applies the cell's filter row and writes it (consecutive threads,
// consecutive addresses).  Ua may alias U_out, U0 may alias U_out (the cells of a
// pass are read completely before any of them is written).  With a state offset
// the three states are offsets from the cell's reference: the stage combination
// is the same (c0 + c1 = 1), the positivity test sees the absolute state, a
// projected component is stored back as its offset, and the filter (which keeps
// constants) acts on the offset.
template <int NS, typename real>
__global__ void __launch_bounds__(kThreads) k_rk_stage(const real* U0, const real* Ua, const real* __restrict__ R,
                                                       const real* __restrict__ dt, real* U_out, StateOff so,
                                                       const real* __restrict__ filt0, const real* __restrict__ filt1,
                                                       int64_t n_cells, int64_t n_prism, double c0, double c1,
                                                       double c2, double p_floor) {
    constexpr int NSV = NS * 5, CPB = kThreads / NS;
    __shared__ real s_f[2][NS * NS];
    __shared__ real s_x[CPB * NSV];
    for (int i = threadIdx.x; i < NS * NS; i += kThreads) {
        s_f[0][i] = filt0[i];
        s_f[1][i] = filt1[i];
    }
    __syncthreads();
    const int t = threadIdx.x;
    for (int64_t cb = (int64_t)blockIdx.x * CPB; cb < n_cells; cb += (int64_t)gridDim.x * CPB) {
        const int nc = (int)(n_cells - cb < CPB ? n_cells - cb : CPB);
        if (t < nc * NS) {
            const int64_t cell = cb + t / NS;
            const int64_t i = cb * NS + t;
            const double d = (double)dt[i];
            double q[5], qa[5], off[5];
            so.row(cell, off);
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                double val = c0 * (double)U0[i * 5 + v] + c2 * d * (double)R[i * 5 + v];
                if (c1 != 0.0) val += c1 * (double)Ua[i * 5 + v];
                q[v] = (double)(real)val;  // stage value in the storage precision
                qa[v] = so.mode ? q[v] + off[v] : q[v];
            }
            const bool low_rho = !(qa[0] > 1e-6);
            const double rho = low_rho ? 1e-6 : qa[0];
            double vx = qa[1] / rho, vy = qa[2] / rho, vz = qa[3] / rho;
            const double vm = sqrt(vx * vx + vy * vy + vz * vz);
            const bool fast = vm > 1e4;
            double m1 = qa[1], m2 = qa[2], m3 = qa[3];
            if (fast) {
                const double f = 1e4 / vm;
                vx *= f; vy *= f; vz *= f;
                m1 = rho * vx; m2 = rho * vy; m3 = rho * vz;
            }
            const double ke = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
            double E = qa[4];
            const bool low_p = 0.4 * (E - ke) < p_floor;
            if (low_p) E = p_floor / 0.4 + ke;
            // a component the projection left alone is stored as it was (exact); a projected one
            // as its (absolute) value less the offset
            real* x = s_x + t * 5;
            x[0] = low_rho ? (real)(rho - off[0]) : (real)q[0];
            x[1] = fast ? (real)(m1 - off[1]) : (real)q[1];
            x[2] = fast ? (real)(m2 - off[2]) : (real)q[2];
            x[3] = fast ? (real)(m3 - off[3]) : (real)q[3];
            x[4] = low_p ? (real)(E - off[4]) : (real)q[4];
        }
        __syncthreads();
        for (int o = t; o < nc * NSV; o += kThreads) {
            const int cc = o / NSV, r = o % NSV, s = r / 5, v = r % 5;
            const real* F = s_f[cb + cc < n_prism ? 0 : 1];
            const real* x = s_x + cc * NSV;
            real acc = 0;
            for (int j = 0; j < NS; ++j) acc += F[s * NS + j] * x[j * 5 + v];
            U_out[cb * NSV + o] = acc;
        }
        __syncthreads();
    }
}

// partial sums of x^2 in FP64, one per workgroup, over the logical array x[a0, a0 + n0) followed
// by x[a1, a1 + n1) (two ranges: the prism and the tet cells of one case of a batched mesh; the
// same partition and order as one contiguous array of n0 + n1 values)
template <typename real>
__global__ void __launch_bounds__(kThreads) k_sq_partial(const real* __restrict__ x, int64_t a0, int64_t n0,
                                                         int64_t a1, int64_t n1, double* part) {
    __shared__ double s[kThreads];
    double acc = 0.0;
    const int64_t n = n0 + n1;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const double v = (double)x[i < n0 ? a0 + i : a1 + (i - n0)];
        acc += v * v;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int m = kThreads / 2; m > 0; m >>= 1) {
        if (threadIdx.x < m) s[threadIdx.x] += s[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kThreads) k_sum_partials(const double* part, int n_part, double n_total, double* out) {
    __shared__ double s[kThreads];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n_part; i += kThreads) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int m = kThreads / 2; m > 0; m >>= 1) {
        if (threadIdx.x < m) s[threadIdx.x] += s[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = n_total > 0 ? sqrt(s[0] / n_total) : 0.0;
}

static int grid_for(int64_t n, int per) {
    int64_t b = (n + per - 1) / per;
    return (int)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

template <typename real>
int launch_dt_local(const real* U, StateOff so, const real* mu_t, const real* det, const real* mfs,
                    const real* vol, real* dt, int64_t n_cells, int ns, double mu, int order, int nut_mode,
                    hipStream_t st) {
#define AFR_DT_CELLS(NSV)                                                                                    \
    hipLaunchKernelGGL((k_dt_cells<NSV, real>), dim3(grid_for(n_cells, kThreads / NSV)), dim3(kThreads), 0, st, U, \
                       so, mu_t, det, mfs, vol, dt, n_cells, mu, order, nut_mode)
    if (ns == 1)
        AFR_DT_CELLS(1);
    else if (ns == 8)
        AFR_DT_CELLS(8);
    else if (ns == 27)
        AFR_DT_CELLS(27);
    else if (ns == 64)
        AFR_DT_CELLS(64);
    else
        hipLaunchKernelGGL((k_dt_local<real>), dim3(grid_for(n_cells * ns, kThreads)), dim3(kThreads), 0, st, U,
                           so, mu_t, det, mfs, vol, dt, n_cells, ns, mu, order, nut_mode);
#undef AFR_DT_CELLS
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename real>
int launch_rk_stage(int ns, const real* U0, const real* Ua, const real* R, const real* dt, real* U_out,
                    StateOff so, const real* filt0, const real* filt1, int64_t n_cells, int64_t n_prism,
                    double c0, double c1, double c2, double p_floor, hipStream_t st) {
    const int grid = grid_for(n_cells, ns > 0 ? kThreads / ns : 1);
    if (ns == 8)
        hipLaunchKernelGGL((k_rk_stage<8, real>), dim3(grid), dim3(kThreads), 0, st, U0, Ua, R, dt, U_out, so,
                           filt0, filt1, n_cells, n_prism, c0, c1, c2, p_floor);
    else if (ns == 27)
        hipLaunchKernelGGL((k_rk_stage<27, real>), dim3(grid), dim3(kThreads), 0, st, U0, Ua, R, dt, U_out, so,
                           filt0, filt1, n_cells, n_prism, c0, c1, c2, p_floor);
    else if (ns == 64)
        hipLaunchKernelGGL((k_rk_stage<64, real>), dim3(grid), dim3(kThreads), 0, st, U0, Ua, R, dt, U_out, so,
                           filt0, filt1, n_cells, n_prism, c0, c1, c2, p_floor);
    else
        return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename real>
__global__ void __launch_bounds__(kThreads) k_fill(real* __restrict__ x, int64_t n, double v) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        x[i] = (real)v;
}

template <typename real>
int launch_fill(real* x, int64_t n, double v, hipStream_t st) {
    hipLaunchKernelGGL((k_fill<real>), dim3(grid_for(n, kThreads)), dim3(kThreads), 0, st, x, n, v);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template int launch_fill<float>(float*, int64_t, double, hipStream_t);
template int launch_fill<double>(double*, int64_t, double, hipStream_t);

template <typename real>
int launch_norm(const real* x, int64_t n, double n_total, double* part, int n_part_max, double* out, hipStream_t st) {
    return launch_norm2(x, 0, n, 0, 0, n_total, part, n_part_max, out, st);
}

template <typename real>
int launch_norm2(const real* x, int64_t a0, int64_t n0, int64_t a1, int64_t n1, double n_total, double* part,
                 int n_part_max, double* out, hipStream_t st) {
    int g = grid_for(n0 + n1, kThreads * 8);
    g = g < n_part_max ? g : n_part_max;
    hipLaunchKernelGGL((k_sq_partial<real>), dim3(g), dim3(kThreads), 0, st, x, a0, n0, a1, n1, part);
    if (hipGetLastError() != hipSuccess) return -1;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kThreads), 0, st, (const double*)part, g, n_total, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template int launch_dt_local<float>(const float*, StateOff, const float*, const float*, const float*, const float*,
                                    float*, int64_t, int, double, int, int, hipStream_t);
template int launch_dt_local<double>(const double*, StateOff, const double*, const double*, const double*,
                                     const double*, double*, int64_t, int, double, int, int, hipStream_t);
template int launch_rk_stage<float>(int, const float*, const float*, const float*, const float*, float*, StateOff,
                                    const float*, const float*, int64_t, int64_t, double, double, double, double,
                                    hipStream_t);
template int launch_rk_stage<double>(int, const double*, const double*, const double*, const double*, double*,
                                     StateOff, const double*, const double*, int64_t, int64_t, double, double, double,
                                     double, hipStream_t);
template int launch_norm2<float>(const float*, int64_t, int64_t, int64_t, int64_t, double, double*, int, double*,
                                 hipStream_t);
template int launch_norm2<double>(const double*, int64_t, int64_t, int64_t, int64_t, double, double*, int, double*,
                                  hipStream_t);
template int launch_norm<float>(const float*, int64_t, double, double*, int, double*, hipStream_t);
template int launch_norm<double>(const double*, int64_t, double, double*, int, double*, hipStream_t);

}  // namespace afr_step
