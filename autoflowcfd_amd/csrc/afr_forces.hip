// FP64 wall force integral (pressure and viscous traction) over the owner-primary
// WALL face records, as src/autoflowcfd/postprocess/fr_coefficients.py:27-152:
// p, grad u (gradient of the primitive velocity, state floors of
// src/autoflowcfd/core/fr_solver/state.py:76-97) and mu_t extrapolated from the
// owner cell to the record's flux points, unit record normal, weights = record area
// x normalised tensor Gauss weights.  One wavefront per record, FP64 throughout,
// deterministic two-pass reduction.
#include <hip/hip_runtime.h>

#include "../../include/afr.h"
#include "afr_forces.hpp"

namespace afr_forces {

constexpr int kWaves = 4;

template <int P, typename real>
__global__ void __launch_bounds__(64 * kWaves) k_wall_forces(ForceArgs<real> a) {
    constexpr int N1 = P + 1, NS = N1 * N1 * N1, NFP = N1 * N1;
    __shared__ double s_q[kWaves][NS * 4];     // u, v, w, p
    __shared__ double s_g[kWaves][NS * 9];     // grad u
    __shared__ double s_f[kWaves][NFP * 6];
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), t = threadIdx.x & 63;
    double* q = s_q[wid];
    double* gd = s_g[wid];
    double* fo = s_f[wid];
    double wgt[N1];
    {
        // normalised tensor Gauss-Legendre weights of the face grid
        const double* w1 = a.w1d;
        double sum = 0.0;
        for (int i = 0; i < N1; ++i) wgt[i] = w1[i];
        for (int i = 0; i < N1; ++i)
            for (int j = 0; j < N1; ++j) sum += wgt[i] * wgt[j];
        for (int i = 0; i < N1; ++i) wgt[i] /= sqrt(sum);
    }
    for (int64_t r = (int64_t)blockIdx.x * kWaves + wid; r < a.n_wall; r += (int64_t)gridDim.x * kWaves) {
        const int64_t c = a.wrec_owner[r];
        const int slot = a.wrec_oslot[r];
        const bool is_prism = c < a.n_prism;
        const real* D = is_prism ? a.D0 : a.D1;
        const real* E = (is_prism ? a.E0 : a.E1) + slot * NFP * NS;
        if (t < NS) {
            const int64_t i = c * NS + t;
            double u5[5];
            afr::load_state(a.U, i, a.so, c, u5);
            const double rho = fmax(u5[0], 1e-10);
            const double u = u5[1] / rho, v = u5[2] / rho, w = u5[3] / rho;
            const double ke = 0.5 * (u * u + v * v + w * w);
            q[t * 4 + 0] = u;
            q[t * 4 + 1] = v;
            q[t * 4 + 2] = w;
            q[t * 4 + 3] = fmax((u5[4] - rho * ke) * 0.4, 1.0);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int o = t; o < 3 * NS; o += 64) {
            const int s = o / 3, f = o % 3;
            double d0 = 0, d1 = 0, d2 = 0;
            for (int j = 0; j < NS; ++j) {
                const double x = q[j * 4 + f];
                d0 += (double)D[(0 * NS + s) * NS + j] * x;
                d1 += (double)D[(1 * NS + s) * NS + j] * x;
                d2 += (double)D[(2 * NS + s) * NS + j] * x;
            }
            const real* ad = a.adj + (c * NS + s) * 9;
            const double idet = 1.0 / (double)a.det[c * NS + s];
            for (int n = 0; n < 3; ++n)
                gd[s * 9 + f * 3 + n] =
                    ((double)ad[0 * 3 + n] * d0 + (double)ad[1 * 3 + n] * d1 + (double)ad[2 * 3 + n] * d2) * idet;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (t < NFP) {
            double p = 0.0, gv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, mt = 0.0;
            for (int s = 0; s < NS; ++s) {
                const double e = (double)E[t * NS + s];
                p += e * q[s * 4 + 3];
                for (int k = 0; k < 9; ++k) gv[k] += e * gd[s * 9 + k];
                if (a.mu_t) mt += e * (double)a.mu_t[c * NS + s];
            }
            const double n0 = a.wrec_geom[r * 4 + 0], n1 = a.wrec_geom[r * 4 + 1], n2 = a.wrec_geom[r * 4 + 2];
            const double aw = a.wrec_geom[r * 4 + 3] * wgt[t / N1] * wgt[t % N1];
            const double m = a.mu + mt;
            const double div = gv[0] + gv[4] + gv[8];
            const double lam = -2.0 / 3.0 * m * div;
            double tau[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) tau[i][j] = m * (gv[i * 3 + j] + gv[j * 3 + i]) + (i == j ? lam : 0.0);
            const double nn[3] = {n0, n1, n2};
            for (int i = 0; i < 3; ++i) {
                fo[t * 6 + i] = p * nn[i] * aw;
                fo[t * 6 + 3 + i] = -(tau[i][0] * n0 + tau[i][1] * n1 + tau[i][2] * n2) * aw;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (t < 6) {
            double acc = 0.0;
            for (int k = 0; k < NFP; ++k) acc += fo[k * 6 + t];
            a.part[r * 6 + t] = acc;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

__global__ void __launch_bounds__(256) k_sum_forces(const double* part, int64_t n, double* out) {
    __shared__ double s[256];
    for (int comp = 0; comp < 6; ++comp) {
        double acc = 0.0;
        for (int64_t i = threadIdx.x; i < n; i += 256) acc += part[i * 6 + comp];
        s[threadIdx.x] = acc;
        __syncthreads();
        for (int m = 128; m > 0; m >>= 1) {
            if (threadIdx.x < m) s[threadIdx.x] += s[threadIdx.x + m];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[comp] = s[0];
        __syncthreads();
    }
}

int launch_sum_forces(const double* part, int64_t n, double* out, hipStream_t st);
template <int P, typename real>
int launch_wall_forces_records(const ForceArgs<real>& a, hipStream_t st);

template <int P, typename real>
int launch_wall_forces(const ForceArgs<real>& a, hipStream_t st) {
    if (launch_wall_forces_records<P, real>(a, st)) return -1;
    return launch_sum_forces(a.part, a.n_wall, a.out, st);
}

int launch_sum_forces(const double* part, int64_t n, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_forces, dim3(1), dim3(256), 0, st, part, n, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int P, typename real>
int launch_wall_forces_records(const ForceArgs<real>& a, hipStream_t st) {
    if (a.n_wall <= 0) return 0;
    const int64_t need = (a.n_wall + kWaves - 1) / kWaves;
    hipLaunchKernelGGL((k_wall_forces<P, real>), dim3((unsigned)(need < 4096 ? need : 4096)), dim3(64 * kWaves), 0,
                       st, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template int launch_wall_forces_records<1, float>(const ForceArgs<float>&, hipStream_t);
template int launch_wall_forces_records<2, float>(const ForceArgs<float>&, hipStream_t);
template int launch_wall_forces_records<1, double>(const ForceArgs<double>&, hipStream_t);
template int launch_wall_forces_records<2, double>(const ForceArgs<double>&, hipStream_t);
template int launch_wall_forces<1, float>(const ForceArgs<float>&, hipStream_t);
template int launch_wall_forces<2, float>(const ForceArgs<float>&, hipStream_t);
template int launch_wall_forces<1, double>(const ForceArgs<double>&, hipStream_t);
template int launch_wall_forces<2, double>(const ForceArgs<double>&, hipStream_t);
// P3 (tet meshes, FP64 storage)
template int launch_wall_forces_records<3, double>(const ForceArgs<double>&, hipStream_t);
template int launch_wall_forces<3, double>(const ForceArgs<double>&, hipStream_t);

}  // namespace afr_forces
