// SST k-omega point functions shared by the SST kernels and the P0 phase of the order
// continuation: the model constants and the F1 / F2 blend functions of
// src/autoflowcfd/core/turbulence/sst.py:21-120, in FP64 and in FP32 (FP32 storage path).
#pragma once
#include <hip/hip_runtime.h>

namespace afr_sst {

constexpr double SK1 = 0.85, SK2 = 1.0, SW1 = 0.5, SW2 = 0.856, B1 = 0.075, B2 = 0.0828, A1 = 0.31, KAP = 0.41,
                 BSTAR = 0.09;

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double finite_or(double x, double alt) { return isfinite(x) ? x : alt; }

__device__ inline double blend_F1(double k, double w, double d, double nu, double rho, double CD) {
    w = dmax(w, 1e-10);
    k = dmax(k, 1e-10);
    d = dmax(d, 1e-10);
    const double t1 = sqrt(k) / (BSTAR * w * d);
    const double t2 = 500.0 * nu / (d * d * w);
    const double t3 = 4.0 * rho * SW2 * k / (dmax(CD, 1e-10) * d * d);
    double a = dmin(dmax(t1, t2), t3);
    a = isfinite(a) ? a : 1e75;
    a = dmin(a, 1e75);
    const double a2 = a * a;
    return tanh(a2 * a2);
}

__device__ inline double blend_F2(double k, double w, double d, double nu) {
    w = dmax(w, 1e-10);
    k = dmax(k, 1e-10);
    d = dmax(d, 1e-10);
    double a = dmax(2.0 * sqrt(k) / (BSTAR * w * d), 500.0 * nu / (d * d * w));
    a = isfinite(a) ? a : 1e150;
    a = dmin(a, 1e150);
    return tanh(a * a);
}

// Blend functions in FP32 (FP32 storage path): tanh of the same clipped arguments; an infinite
// or non-finite argument saturates tanh at 1 exactly as the FP64 form's 1e75 / 1e150 caps do.
__device__ inline float blend_F1_f(float k, float w, float d, float nu, float rho, float CD) {
    w = fmaxf(w, 1e-10f);
    k = fmaxf(k, 1e-10f);
    d = fmaxf(d, 1e-10f);
    const float t1 = sqrtf(k) / (float(BSTAR) * w * d);
    const float t2 = 500.0f * nu / (d * d * w);
    const float t3 = 4.0f * rho * float(SW2) * k / (fmaxf(CD, 1e-10f) * d * d);
    float a = fminf(fmaxf(t1, t2), t3);
    a = isfinite(a) ? a : 1e30f;
    a = fminf(a, 1e30f);
    const float a2 = a * a;
    return tanhf(a2 * a2);
}

__device__ inline float blend_F2_f(float k, float w, float d, float nu) {
    w = fmaxf(w, 1e-10f);
    k = fmaxf(k, 1e-10f);
    d = fmaxf(d, 1e-10f);
    float a = fmaxf(2.0f * sqrtf(k) / (float(BSTAR) * w * d), 500.0f * nu / (d * d * w));
    a = isfinite(a) ? a : 1e30f;
    a = fminf(a, 1e30f);
    return tanhf(a * a);
}

template <typename real> struct Blend;
template <> struct Blend<double> {
    static __device__ double F1(double k, double w, double d, double nu, double rho, double CD) {
        return blend_F1(k, w, d, nu, rho, CD);
    }
    static __device__ double F2(double k, double w, double d, double nu) { return blend_F2(k, w, d, nu); }
};
template <> struct Blend<float> {
    static __device__ double F1(double k, double w, double d, double nu, double rho, double CD) {
        return (double)blend_F1_f((float)dmin(k, 1e38), (float)dmin(w, 1e38), (float)dmin(d, 1e38),
                                  (float)nu, (float)rho, (float)dmin(CD, 1e38));
    }
    static __device__ double F2(double k, double w, double d, double nu) {
        return (double)blend_F2_f((float)dmin(k, 1e38), (float)dmin(w, 1e38), (float)dmin(d, 1e38), (float)nu);
    }
};

}  // namespace afr_sst
