// Residual outlier zeroing of one cell (device side), shared by every kernel that applies it.
//
// src/autoflowcfd/core/fr_operators/troubled_cell.py:339-408: for each variable v of a cell,
//   ref = max(median_s |R[s][v]|, 1e-9 mean_s |U[s][v]|, 1e-300),  R[s][v] = 0 where |R| > 1e4 ref,
// with the median of _median_abs_over_sps_kernel (:347-379): the |R| of the cell are sorted (NaN
// last, as numpy sorts) and the middle value, or the mean of the two middle values for an even
// count, is taken.  Consequences kept here: a few NaN leave a finite median; a median position that
// falls on NaN makes the median NaN, and np.maximum propagates it, so nothing of that variable is
// zeroed; NaN residuals are never zeroed; a NaN field mean zeroes nothing either.
#pragma once
#include <hip/hip_runtime.h>

namespace afr {

// max that propagates NaN from either side, as np.maximum does
template <typename real>
__device__ __forceinline__ real nan_max(real a, real b) { return (a > b || a != a) ? a : b; }

// One wavefront owns one cell whose residual block sR[s * NV + v] (NS points, NV <= 8 variables)
// sits in LDS; lane t = threadIdx.x & 63.  umean(v) returns mean_s |U[s][v]| (only called when the
// exact median is needed).  A certificate runs first: with M = max_s |R[s][v]| and
// tt = 1.0001e-4 M, if at most (NS-1)/2 values are below tt then the median is >= tt or NaN, so
// no value can exceed 1e4 ref and the variable is left as is.  Only when the certificate fails
// for some variable is the median selected exactly, by rank counting (NaN compares false, so it
// takes no rank: the finite values fill the first positions of the sorted order).
template <int NS, int NV, typename real, typename UMean>
__device__ __forceinline__ void outlier_zeroing(real* sR, UMean umean, int t) {
    static_assert(NV >= 1 && NV <= 8, "8 certificate lanes per variable");
    constexpr int NSV = NS * NV, NR = (NSV + 63) / 64;
    constexpr int k1 = (NS - 1) / 2, k2 = NS / 2;
    constexpr int G = 8, J = (NS + G - 1) / G;  // 8 lanes per variable, J values each
    const int v = t >> 3, g = t & 7;
    real mx = 0;
    if (t < NV * G) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int s = g + G * j;
            if (s < NS) {
                const real x = fabs(sR[s * NV + v]);
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
    if (t < NV * G) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int s = g + G * j;
            if (s < NS) cnt += fabs(sR[s * NV + v]) < tt;
        }
    }
#pragma unroll
    for (int m = 1; m < G; m <<= 1) cnt += __shfl_xor(cnt, m);
    if (!__any((t < NV * G) && cnt > k1)) return;
    // exact selection: rank of every value among the finite values of its variable
    real x[NR], f1[NR], f2[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int o = t + 64 * k;
        x[k] = f1[k] = f2[k] = real(0);
        if (o < NSV) {
            const int vv = o % NV;
            x[k] = fabs(sR[o]);
            int lt = 0, le = 0;
#pragma unroll 1
            for (int s2 = 0; s2 < NS; ++s2) {  // rare path: rolled, off the register budget
                const real y = fabs(sR[s2 * NV + vv]);
                lt += y < x[k];
                le += y <= x[k];
            }
            f1[k] = (lt <= k1 && k1 < le) ? real(1) : real(0);
            f2[k] = (lt <= k2 && k2 < le) ? real(1) : real(0);
        }
    }
    // the value at a median position is held by the lanes flagged above (all equal): gather it
    // by a max over the flagged values; -1 when no finite value holds the position (NaN median)
    real med1[NV], med2[NV];
#pragma unroll
    for (int vv = 0; vv < NV; ++vv) {
        real a1 = real(-1), a2 = real(-1);
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const int o = t + 64 * k;
            if (o < NSV && o % NV == vv) {
                if (f1[k] > real(0)) a1 = x[k] > a1 ? x[k] : a1;
                if (f2[k] > real(0)) a2 = x[k] > a2 ? x[k] : a2;
            }
        }
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            const real o1 = __shfl_xor(a1, m), o2 = __shfl_xor(a2, m);
            a1 = o1 > a1 ? o1 : a1;
            a2 = o2 > a2 ? o2 : a2;
        }
        med1[vv] = a1;
        med2[vv] = a2;
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int o = t + 64 * k;
        if (o < NSV) {
            const int vv = o % NV;
            real m1v = med1[0], m2v = med2[0];
#pragma unroll
            for (int q = 1; q < NV; ++q)
                if (vv == q) { m1v = med1[q]; m2v = med2[q]; }
            if (m1v < real(0) || ((NS % 2) == 0 && m2v < real(0))) continue;  // NaN median
            const real med = (NS % 2) ? m1v : real(0.5) * (m1v + m2v);
            real ref = nan_max(med, real(1e-9) * umean(vv));
            ref = nan_max(ref, Tiny<real>::v);
            if (x[k] > real(1e4) * ref) sR[o] = real(0);
        }
    }
}

}  // namespace afr
