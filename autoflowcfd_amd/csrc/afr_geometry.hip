// FP64 metric precompute on the GPU (gfx950).
//
// From vertex coordinates, per cell:
//   det(J), adj(J) at the solution points and adj(J) at the over-integration
//   points (exact Jacobians of straight-sided prisms/tets,
//   src/autoflowcfd/grid/curved_mapping/curved_mapping_exact_jacobian.py:39-119,
//   cofactor form of curved_mapping.py:39-111), GL-quadrature volumes
//   (src/autoflowcfd/grid/high_order/high_order_mesh.py:335-354), the metric
//   flux scale sum_m |adj row m| (src/autoflowcfd/core/fr_solver/solver_geometry.py:18-32),
//   the constant-state divergence G[s][i] = sum_{m,j} Op[s][m,j] adj_fine[j][m][i]
//   used by the perturbation form of the residual, and per face slot the
//   extrapolated adjugate rows E adj[:, axis, :] with the alignment fallback
//   bit of src/autoflowcfd/core/fr_residual/inviscid_kernel.py:148-152.
// Everything is computed in FP64 and rounded once to the storage precision.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/afr.h"

extern void afr_set_error(const std::string& msg);

namespace {

constexpr int kT = 256;

template <int P> struct G {
    static constexpr int N1 = P + 1;
    static constexpr int NS = N1 * N1 * N1;
    static constexpr int NFP = N1 * N1;
    static constexpr int OV = (2 * P < 3 ? 2 * P : 3);
    static constexpr int NF = (OV + 1) * (OV + 1) * (OV + 1);
    static constexpr int CPB = kT / NF;
};

__device__ inline void jac_prism(const double* p, double a, double b, double c, double J[3][3]) {
    double r = (1.0 + a) * (1.0 - b) / 2.0 - 1.0, s = b;
    double l1 = -(r + s) / 2.0, l2 = (1.0 + r) / 2.0, l3 = (1.0 + s) / 2.0;
    double cm = 0.5 * (1.0 - c), cp = 0.5 * (1.0 + c), bb = (1.0 - b) / 4.0;
    double am = -(1.0 - a) / 4.0, ap = -(1.0 + a) / 4.0;
    for (int d = 0; d < 3; ++d) {
        double p0 = p[0 * 3 + d], p1 = p[1 * 3 + d], p2 = p[2 * 3 + d];
        double p3 = p[3 * 3 + d], p4 = p[4 * 3 + d], p5 = p[5 * 3 + d];
        J[d][0] = cm * bb * (p1 - p0) + cp * bb * (p4 - p3);
        J[d][1] = cm * (am * p0 + ap * p1 + 0.5 * p2) + cp * (am * p3 + ap * p4 + 0.5 * p5);
        J[d][2] = 0.5 * ((l1 * p3 + l2 * p4 + l3 * p5) - (l1 * p0 + l2 * p1 + l3 * p2));
    }
}

__device__ inline void jac_tet(const double* p, double a, double b, double c, double J[3][3]) {
    double s = (1.0 + b) * (1.0 - c) / 2.0 - 1.0, t = c;
    for (int d = 0; d < 3; ++d) {
        double e1 = (p[3 + d] - p[d]) / 2.0, e2 = (p[6 + d] - p[d]) / 2.0, e3 = (p[9 + d] - p[d]) / 2.0;
        J[d][0] = -(s + t) / 2.0 * e1;
        J[d][1] = -(1.0 + a) * (1.0 - c) / 4.0 * e1 + (1.0 - c) / 2.0 * e2;
        J[d][2] = -(1.0 + a) * (1.0 - b) / 4.0 * e1 - (1.0 + b) / 2.0 * e2 + e3;
    }
}

// physical position of reference point (a, b, c) (collapsed prism / tet maps)
__device__ inline void map_prism(const double* p, double a, double b, double c, double x[3]) {
    const double r = (1.0 + a) * (1.0 - b) / 2.0 - 1.0, s = b;
    const double l1 = -(r + s) / 2.0, l2 = (1.0 + r) / 2.0, l3 = (1.0 + s) / 2.0;
    for (int d = 0; d < 3; ++d) {
        const double bot = l1 * p[0 * 3 + d] + l2 * p[1 * 3 + d] + l3 * p[2 * 3 + d];
        const double top = l1 * p[3 * 3 + d] + l2 * p[4 * 3 + d] + l3 * p[5 * 3 + d];
        x[d] = 0.5 * (1.0 - c) * bot + 0.5 * (1.0 + c) * top;
    }
}

__device__ inline void map_tet(const double* p, double a, double b, double c, double x[3]) {
    const double t = c, s = (1.0 + b) * (1.0 - c) / 2.0 - 1.0, r = -(1.0 + a) * (s + t) / 2.0 - 1.0;
    for (int d = 0; d < 3; ++d)
        x[d] = -(1.0 + r + s + t) / 2.0 * p[d] + (1.0 + r) / 2.0 * p[3 + d] + (1.0 + s) / 2.0 * p[6 + d] +
               (1.0 + t) / 2.0 * p[9 + d];
}

// adj[m][i] = det * (J^{-1})[m][i]
__device__ inline double det_adj(const double J[3][3], double A[9]) {
    A[0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    A[3] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]);
    A[6] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    A[1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]);
    A[4] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    A[7] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]);
    A[2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    A[5] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]);
    A[8] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    return J[0][0] * A[0] + J[0][1] * A[3] + J[0][2] * A[6];
}

template <typename real>
struct GeoArgs {
    int64_t begin, n;
    const double* nodes;
    const int* conn;
    int nv;  // vertices per cell
    const double* sps;
    const double* fine;
    const double* w;
    const double* op;  // (NS, 3*NF) FP64
    const double* E;   // (5, NFP, NS) FP64, slot order
    const double* tn;  // (C, 5, 3) FP64
    real* det;
    real* adj;
    real* adjf;
    real* gcl;
    double* vol;
    double* det_mean;
    real* mfs;
    real* adjrow;
    int* fallback;
    int* bad;
    double* xyz;  // (C, NS, 3) FP64 solution-point positions, may be null
    int nslot;
};

template <int P, int CT, typename real>
__global__ void __launch_bounds__(kT) k_metrics(GeoArgs<real> g) {
    using O = G<P>;
    constexpr int NS = O::NS, NF = O::NF, NFP = O::NFP, CPB = O::CPB;
    __shared__ double s_af[CPB * NF * 9];
    __shared__ double s_a[CPB * NS * 9];
    __shared__ double s_d[CPB * NS];
    __shared__ double s_x[CPB * 18];
    const int t = threadIdx.x;
    const int64_t n_tiles = (g.n + CPB - 1) / CPB;
    for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const int64_t first = tile * CPB;  // local index within the type block
        const int ncell = (int)((g.n - first) < CPB ? (g.n - first) : CPB);
        if (t < CPB * g.nv * 3) {
            const int c = t / (g.nv * 3), r = t % (g.nv * 3), k = r / 3, d = r % 3;
            if (c < ncell) {
                const int node = g.conn[(first + c) * g.nv + k];
                s_x[c * 18 + k * 3 + d] = g.nodes[(int64_t)node * 3 + d];
            }
        }
        __syncthreads();
        // fine points
        if (t < CPB * NF) {
            const int c = t / NF, j = t % NF;
            if (c < ncell) {
                double J[3][3], A[9];
                const double* x = g.fine + j * 3;
                if (CT == 0) jac_prism(s_x + c * 18, x[0], x[1], x[2], J);
                else jac_tet(s_x + c * 18, x[0], x[1], x[2], J);
                double dt = det_adj(J, A);
                if (!(dt > 0.0)) atomicOr(g.bad, 1);
                const int64_t gc = g.begin + first + c;
                for (int k = 0; k < 9; ++k) {
                    s_af[(c * NF + j) * 9 + k] = A[k];
                    g.adjf[(gc * NF + j) * 9 + k] = (real)A[k];
                }
            }
        }
        // solution points
        if (t < CPB * NS) {
            const int c = t / NS, s = t % NS;
            if (c < ncell) {
                double J[3][3], A[9];
                const double* x = g.sps + s * 3;
                if (CT == 0) jac_prism(s_x + c * 18, x[0], x[1], x[2], J);
                else jac_tet(s_x + c * 18, x[0], x[1], x[2], J);
                double dt = det_adj(J, A);
                if (!(dt > 0.0)) atomicOr(g.bad, 1);
                const int64_t gc = g.begin + first + c;
                s_d[c * NS + s] = dt;
                g.det[gc * NS + s] = (real)dt;
                double m = 0.0;
                for (int mm = 0; mm < 3; ++mm)
                    m += sqrt(A[mm * 3] * A[mm * 3] + A[mm * 3 + 1] * A[mm * 3 + 1] + A[mm * 3 + 2] * A[mm * 3 + 2]);
                g.mfs[gc * NS + s] = (real)m;
                if (g.xyz) {
                    double X[3];
                    if (CT == 0) map_prism(s_x + c * 18, x[0], x[1], x[2], X);
                    else map_tet(s_x + c * 18, x[0], x[1], x[2], X);
                    for (int d = 0; d < 3; ++d) g.xyz[(gc * NS + s) * 3 + d] = X[d];
                }
                for (int k = 0; k < 9; ++k) {
                    s_a[(c * NS + s) * 9 + k] = A[k];
                    g.adj[(gc * NS + s) * 9 + k] = (real)A[k];
                }
            }
        }
        __syncthreads();
        // constant-state divergence (GCL defect of the discrete operator)
        for (int o = t; o < CPB * NS * 3; o += kT) {
            const int c = o / (NS * 3), r = o % (NS * 3), s = r / 3, i = r % 3;
            if (c >= ncell) continue;
            double acc = 0.0;
            for (int m = 0; m < 3; ++m)
                for (int j = 0; j < NF; ++j)
                    acc += g.op[s * 3 * NF + m * NF + j] * s_af[(c * NF + j) * 9 + m * 3 + i];
            g.gcl[((g.begin + first + c) * NS + s) * 3 + i] = (real)acc;
        }
        // volumes, mean det
        if (t < CPB) {
            if (t < ncell) {
                double v = 0.0, dm = 0.0;
                for (int s = 0; s < NS; ++s) {
                    v += g.w[s] * s_d[t * NS + s];
                    dm += s_d[t * NS + s];
                }
                g.vol[g.begin + first + t] = v;
                g.det_mean[g.begin + first + t] = dm / NS;
            }
        }
        // extrapolated adjugate rows per slot and alignment fallback
        if (t < CPB * 5) {
            const int c = t / 5, slot = t % 5;
            if (c < ncell) {
                const int64_t gc = g.begin + first + c;
                int bits = 0;
                if (slot < g.nslot) {
                    const int code = slot < 3 ? slot : (slot == 3 ? 4 : 5);
                    const int axis = code >> 1;
                    const double side = (code & 1) ? 1.0 : -1.0;
                    const double* tn = g.tn + (gc * 5 + slot) * 3;
                    for (int fp = 0; fp < NFP; ++fp) {
                        double ar[3] = {0, 0, 0};
                        for (int s = 0; s < NS; ++s) {
                            const double e = g.E[(slot * NFP + fp) * NS + s];
                            const double* A = s_a + (c * NS + s) * 9 + axis * 3;
                            ar[0] += e * A[0];
                            ar[1] += e * A[1];
                            ar[2] += e * A[2];
                        }
                        double mag = sqrt(ar[0] * ar[0] + ar[1] * ar[1] + ar[2] * ar[2]);
                        mag = mag > 1e-300 ? mag : 1e-300;
                        double al = (ar[0] * tn[0] + ar[1] * tn[1] + ar[2] * tn[2]) / mag * side;
                        if (al < 0.5) bits |= (1 << fp);
                        for (int i = 0; i < 3; ++i) g.adjrow[((gc * 5 + slot) * NFP + fp) * 3 + i] = (real)ar[i];
                    }
                } else {
                    for (int k = 0; k < NFP * 3; ++k) g.adjrow[(gc * 5 + slot) * NFP * 3 + k] = (real)0;
                }
                g.fallback[gc * 5 + slot] = bits;
            }
        }
        __syncthreads();
    }
}

template <int P, typename real>
int launch(const AfrGeomDesc* d, hipStream_t st) {
    using O = G<P>;
    for (int ct = 0; ct < 2; ++ct) {
        const int64_t n = ct == 0 ? d->n_prism : d->n_cells - d->n_prism;
        if (n <= 0) continue;
        GeoArgs<real> g;
        g.begin = ct == 0 ? 0 : d->n_prism;
        g.n = n;
        g.nodes = d->nodes;
        g.conn = ct == 0 ? d->prism_conn : d->tet_conn;
        g.nv = ct == 0 ? 6 : 4;
        g.sps = d->sps;
        g.fine = d->fine;
        g.w = d->w_sps;
        g.op = d->op64[ct];
        g.E = d->E64[ct];
        g.tn = d->slot_tn64;
        g.det = (real*)d->det;
        g.adj = (real*)d->adj;
        g.adjf = (real*)d->adj_fine;
        g.gcl = (real*)d->gcl;
        g.vol = d->volume;
        g.det_mean = d->det_mean;
        g.mfs = (real*)d->metric_flux_scale;
        g.adjrow = (real*)d->slot_adjrow;
        g.fallback = d->slot_fallback;
        g.bad = d->bad_det;
        g.xyz = d->sps_xyz;
        g.nslot = ct == 0 ? 5 : 4;
        int64_t tiles = (n + O::CPB - 1) / O::CPB;
        int grid = (int)(tiles < 8192 ? tiles : 8192);
        if (ct == 0)
            hipLaunchKernelGGL((k_metrics<P, 0, real>), dim3(grid), dim3(kT), 0, st, g);
        else
            hipLaunchKernelGGL((k_metrics<P, 1, real>), dim3(grid), dim3(kT), 0, st, g);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            afr_set_error(std::string("afr_metrics launch: ") + hipGetErrorString(e));
            return -1;
        }
    }
    return 0;
}

}  // namespace

extern "C" int afr_metrics(const AfrGeomDesc* d, void* stream) {
    if (!d || d->abi_version != AFR_ABI_VERSION) {
        afr_set_error("afr_metrics: bad descriptor");
        return -1;
    }
    if (d->order < 1 || d->order > 3 || d->n_cells <= 0 || d->n_prism < 0 || d->n_prism > d->n_cells) {
        afr_set_error("afr_metrics: bad order or cell counts");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    if (d->dtype == AFR_F64)
        return d->order == 1 ? launch<1, double>(d, st) : d->order == 2 ? launch<2, double>(d, st) : launch<3, double>(d, st);
    return d->order == 1 ? launch<1, float>(d, st) : d->order == 2 ? launch<2, float>(d, st) : launch<3, float>(d, st);
}
