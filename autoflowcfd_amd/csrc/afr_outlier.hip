// Residual outlier zeroing of a whole field, as one entry point:
//   suppress_residual_outliers  src/autoflowcfd/core/fr_operators/troubled_cell.py:382-408
// The residual kernels apply the same device routine (afr_outlier.hpp) to each cell's block before
// it leaves LDS; this entry point runs it over a field held in device memory, one wavefront per
// cell, for callers that hold a residual of their own (any order, any number of variables).
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/afr.h"
#include "afr_physics.hpp"
#include "afr_outlier.hpp"

using namespace afr;

namespace {

constexpr int kWaves = 4;

// LDS ordering inside one wavefront (LDS operations of a wave execute in program order)
__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NS, int NV, typename real>
__global__ void __launch_bounds__(64 * kWaves) k_suppress(real* __restrict__ R, const real* __restrict__ U,
                                                          int64_t n_cells, int u_stride) {
    __shared__ real s_r[kWaves][NS * NV];
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), t = threadIdx.x & 63;
    real* sR = s_r[wid];
    for (int64_t cell = (int64_t)blockIdx.x * kWaves + wid; cell < n_cells; cell += (int64_t)gridDim.x * kWaves) {
        real* Rc = R + cell * (NS * NV);
        for (int o = t; o < NS * NV; o += 64) sR[o] = Rc[o];
        wsync();
        const real* Uc = U + cell * NS * u_stride;
        outlier_zeroing<NS, NV, real>(sR, [Uc, u_stride](int v) {
            real m = real(0);
            for (int s = 0; s < NS; ++s) m += fabs(Uc[s * u_stride + v]);
            return m / real(NS);
        }, t);
        wsync();
        for (int o = t; o < NS * NV; o += 64) Rc[o] = sR[o];
        wsync();
    }
}

template <int NS, int NV, typename real>
int launch(void* R, const void* U, int64_t n_cells, int u_stride, hipStream_t st) {
    int64_t blocks = (n_cells + kWaves - 1) / kWaves;
    blocks = blocks < 8192 ? blocks : 8192;
    if (blocks < 1) return 0;
    hipLaunchKernelGGL((k_suppress<NS, NV, real>), dim3((unsigned)blocks), dim3(64 * kWaves), 0, st, (real*)R,
                       (const real*)U, n_cells, u_stride);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NS, typename real>
int by_nv(int nv, void* R, const void* U, int64_t n, int us, hipStream_t st) {
    switch (nv) {
        case 1: return launch<NS, 1, real>(R, U, n, us, st);
        case 2: return launch<NS, 2, real>(R, U, n, us, st);
        case 5: return launch<NS, 5, real>(R, U, n, us, st);
        case 7: return launch<NS, 7, real>(R, U, n, us, st);
        default: return -2;
    }
}

template <typename real>
int by_ns(int ns, int nv, void* R, const void* U, int64_t n, int us, hipStream_t st) {
    switch (ns) {
        case 8: return by_nv<8, real>(nv, R, U, n, us, st);
        case 27: return by_nv<27, real>(nv, R, U, n, us, st);
        case 64: return by_nv<64, real>(nv, R, U, n, us, st);
        default: return -2;
    }
}

}  // namespace

void afr_set_error(const std::string& msg);

extern "C" int afr_suppress_residual_outliers(void* R, const void* U, int64_t n_cells, int32_t ns, int32_t nv,
                                              int32_t u_stride, int32_t dtype, void* stream) {
    if (!R || !U || n_cells < 0 || nv < 1 || u_stride < nv || (dtype != AFR_F32 && dtype != AFR_F64)) {
        afr_set_error("afr_suppress_residual_outliers: bad argument");
        return -1;
    }
    if (ns == 1) return 0;  // one point per cell: |R| > 1e4 max(|R|, ...) never holds
    hipStream_t st = (hipStream_t)stream;
    const int rc = dtype == AFR_F64 ? by_ns<double>(ns, nv, R, U, n_cells, u_stride, st)
                                    : by_ns<float>(ns, nv, R, U, n_cells, u_stride, st);
    if (rc == -2) {
        afr_set_error("afr_suppress_residual_outliers: supported ns 1, 8, 27, 64 and nv 1, 2, 5, 7");
        return -1;
    }
    if (rc) {
        afr_set_error("afr_suppress_residual_outliers: launch failed");
        return -1;
    }
    return 0;
}
