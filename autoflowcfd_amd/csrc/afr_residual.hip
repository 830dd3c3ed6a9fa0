// Cell-centric FR residual kernels for gfx950 and the afr C ABI.
//
// Tets: one workgroup of 256 threads processes tiles of CPB cells at one
// order with the dense operators (prisms use the Kronecker-factored kernel of
// afr_prism.hip).  Per tile, everything stays in LDS:
//   1. primitive states of the tile's cells,
//   2. contravariant fluxes at the over-integration points
//      (Q_fine = c2f Q, F~ = adj_fine F(Q_fine)),
//   3. the FR interface jumps of every face slot (own trace E Q, other side
//      by cross-interpolation rows or a ghost state, AUSM+up),
//   4. residual = -(Op F~ + sum_slots g' jump) / det  with Op = f2c D_fine,
//   5. outlier suppression (median over solution points),
// and the tile is written once.  No scatter, coloring or atomics.
//
// Reference semantics: src/autoflowcfd/core/fr_residual/inviscid.py:139-345,
// src/autoflowcfd/core/fr_residual/inviscid_kernel.py:87-284,
// src/autoflowcfd/core/fr_operators/troubled_cell.py:339-408.
#include <hip/hip_runtime.h>
#include <vector>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "../../include/afr.h"
#include "afr_physics.hpp"
#include "afr_prism.hpp"
#include "afr_viscous.hpp"
#include "afr_step.hpp"
#include "afr_sst.hpp"
#include "afr_forces.hpp"
#include "afr_outlier.hpp"

using namespace afr;

static thread_local std::string g_err;

static int fail(const char* msg) {
    g_err = msg;
    return -1;
}

void afr_set_error(const std::string& msg) { g_err = msg; }

#define HIP_TRY(x)                                                              \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) {                                                 \
            g_err = std::string(#x) + ": " + hipGetErrorString(e_);             \
            return -1;                                                          \
        }                                                                       \
    } while (0)

struct AfrMesh {
    AfrMeshDesc d;
    // host copies of the case offsets of a batched mesh (cell_case set)
    std::vector<int64_t> case_prism_off, case_tet_off, case_wall_off;
    // the state storage the kernels see: a single-case offset copied here at creation (passed by
    // value), a batched mesh's device table, or none
    afr::StateOff so{};
};

namespace {

constexpr int kThreads = 256;
constexpr int kSlots = 5;

template <int P> struct Ord {
    static constexpr int N1 = P + 1;
    static constexpr int NS = N1 * N1 * N1;
    static constexpr int NFP = N1 * N1;
    static constexpr int OV = (2 * P < 3 ? 2 * P : 3);
    static constexpr int NF = (OV + 1) * (OV + 1) * (OV + 1);
    static constexpr int CPB = kThreads / NF;  // cells per tile
};

template <int CT> struct Cell {
    static constexpr int NSLOT = CT == 0 ? 5 : 4;
    __device__ static int code(int slot) {
        // prism slots: a=-1, a=+1, b=-1, c=-1, c=+1 ; tet slots: a=-1, a=+1, b=-1, c=-1
        return slot < 3 ? slot : (slot == 3 ? 4 : 5);
    }
};

template <typename real>
struct Args {
    int64_t cell_begin, n_cells_type;
    const real* Q;        // (C, NS, 5) primitive
    const real* umean;    // (C, 5) mean |U| per cell and variable (outlier reference)
    real* R;
    const real* det;
    const real* adjf;
    const int* slot_nb;
    const int* slot_blk;
    const int* slot_mask;
    const int* slot_bc;
    const int* slot_fb;
    const real* slot_tn;
    const real* slot_adjrow;
    const real* rows;
    const real* c2f;
    const real* op;
    const real* E;
    const real* gL;
    const real* gR;
    const int* fp_of_sp;
    const int* coord_of_sp;
    const real* gcl;      // (C, NS, 3) constant-state divergence
    const double* bcp;    // (8, 5) per-BC reference states ((n_cases, 8, 5) with cell_case)
    const int* cell_case; // null: one case
    double qref[5];       // primitive reference state
    double fref[15];      // F(qref)[i][v]
    int suppress;
    int n_rows;
};

template <typename real>
__global__ void __launch_bounds__(kThreads) k_primitives(const real* __restrict__ U, real* __restrict__ Q, int64_t n,
                                                         Ref5 ref) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        double u[5], q[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) u[v] = (double)U[i * 5 + v];
        cons2prim(u, q);
#pragma unroll
        for (int v = 0; v < 5; ++v) Q[i * 5 + v] = (real)(q[v] - ref.q[v]);
    }
}

// median of |x[0..n)| as troubled_cell.py:347-379 takes it: the sorted values (NaN last), the
// middle one or the mean of the two middle ones for even n; NaN when a median position holds NaN
template <typename real, int N>
__device__ real median_abs(const real* x, int stride) {
    real a[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        real t = x[i * stride];
        a[i] = t < real(0) ? -t : t;
    }
    const int k1 = (N - 1) / 2, k2 = N / 2;
    const real qnan = real(__builtin_nan(""));
    real m1 = qnan, m2 = qnan;
    for (int i = 0; i < N; ++i) {
        int lt = 0, le = 0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            lt += a[j] < a[i];
            le += a[j] <= a[i];
        }
        if (lt <= k1 && k1 < le) m1 = a[i];
        if (lt <= k2 && k2 < le) m2 = a[i];
    }
    return (N % 2) ? m1 : real(0.5) * (m1 + m2);
}

template <int P, int CT, typename real>
__global__ void __launch_bounds__(kThreads) k_inviscid(Args<real> a) {
    using O = Ord<P>;
    constexpr int NS = O::NS, NF = O::NF, NFP = O::NFP, CPB = O::CPB, N1 = O::N1;
    constexpr int NSLOT = Cell<CT>::NSLOT;
    // P3 operator tables (64 x 64 and larger) do not fit LDS beside the tiles: they are read
    // from global memory (L1/L2 resident) instead
    constexpr bool OPS_LDS = P <= 2;
    __shared__ real s_c2f_l[OPS_LDS ? NF * NS : 1];
    __shared__ real s_op_l[OPS_LDS ? NS * 3 * NF : 1];
    __shared__ real s_E_l[OPS_LDS ? NSLOT * NFP * NS : 1];
    const real* const s_c2f = OPS_LDS ? s_c2f_l : a.c2f;
    const real* const s_op = OPS_LDS ? s_op_l : a.op;
    const real* const s_E = OPS_LDS ? s_E_l : a.E;
    __shared__ real s_Q[CPB * NS * 5];
    __shared__ real s_Ft[CPB * 3 * NF * 5];
    __shared__ real s_J[CPB * NSLOT * NFP * 5];
    __shared__ real s_R[CPB * NS * 5];
    __shared__ real s_g[2 * N1];
    __shared__ int s_fp[3 * NS];
    __shared__ int s_co[3 * NS];

    const int t = threadIdx.x;
    if constexpr (OPS_LDS) {
        for (int i = t; i < NF * NS; i += kThreads) s_c2f_l[i] = a.c2f[i];
        for (int i = t; i < NS * 3 * NF; i += kThreads) s_op_l[i] = a.op[i];
        for (int i = t; i < NSLOT * NFP * NS; i += kThreads) s_E_l[i] = a.E[i];
    }
    for (int i = t; i < 3 * NS; i += kThreads) {
        s_fp[i] = a.fp_of_sp[i];
        s_co[i] = a.coord_of_sp[i];
    }
    if (t < N1) {
        s_g[t] = a.gL[t];
        s_g[N1 + t] = a.gR[t];
    }
    __syncthreads();

    const int64_t n_tiles = (a.n_cells_type + CPB - 1) / CPB;
    for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const int64_t first = a.cell_begin + tile * CPB;
        const int ncell = (int)((a.n_cells_type - tile * CPB) < CPB ? (a.n_cells_type - tile * CPB) : CPB);

        // 1. primitive states of the tile (contiguous)
        for (int i = t; i < CPB * NS * 5; i += kThreads)
            s_Q[i] = (i < ncell * NS * 5) ? a.Q[first * NS * 5 + i] : real(0);
        __syncthreads();

        // 2. contravariant fluxes at the fine points
        if (t < CPB * NF) {
            const int c = t / NF, j = t % NF;
            if (c < ncell) {
                real dq[5] = {0, 0, 0, 0, 0};
                for (int s = 0; s < NS; ++s) {
                    const real w = s_c2f[j * NS + s];
#pragma unroll
                    for (int v = 0; v < 5; ++v) dq[v] += w * s_Q[(c * NS + s) * 5 + v];
                }
                // point flux in FP64 about the reference state, then back to storage precision
                double q[5], Fd[3][5];
#pragma unroll
                for (int v = 0; v < 5; ++v) q[v] = a.qref[v] + (double)dq[v];
                euler_flux(q, Fd);
                real F[3][5];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int v = 0; v < 5; ++v) F[i][v] = (real)(Fd[i][v] - a.fref[i * 5 + v]);
                const real* ad = a.adjf + ((first + c) * NF + j) * 9;
                real A[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) A[k] = ad[k];
#pragma unroll
                for (int m = 0; m < 3; ++m)
#pragma unroll
                    for (int v = 0; v < 5; ++v)
                        s_Ft[((c * 3 + m) * NF + j) * 5 + v] = A[m * 3 + 0] * F[0][v] + A[m * 3 + 1] * F[1][v] +
                                                               A[m * 3 + 2] * F[2][v];
            }
        }

        // 3. interface jumps, one thread per (cell, slot, flux point)
        if (t < CPB * NSLOT * NFP) {
            const int c = t / (NSLOT * NFP), r = t % (NSLOT * NFP);
            const int slot = r / NFP, fp = r % NFP;
            if (c < ncell) {
                const int64_t gc = first + c;
                const int64_t sidx = gc * kSlots + slot;
                const int code = Cell<CT>::code(slot);
                const double side = (code & 1) ? 1.0 : -1.0;
                real dqo[5] = {0, 0, 0, 0, 0};
                for (int s = 0; s < NS; ++s) {
                    const real w = s_E[(slot * NFP + fp) * NS + s];
#pragma unroll
                    for (int v = 0; v < 5; ++v) dqo[v] += w * s_Q[(c * NS + s) * 5 + v];
                }
                double qo[5], qn[5];
#pragma unroll
                for (int v = 0; v < 5; ++v) qo[v] = a.qref[v] + (double)dqo[v];
                double tn[3] = {(double)a.slot_tn[sidx * 3 + 0], (double)a.slot_tn[sidx * 3 + 1],
                                (double)a.slot_tn[sidx * 3 + 2]};
                const int bc = a.slot_bc[sidx];
                if (bc == BC_INTERIOR) {
                    real dqn[5] = {0, 0, 0, 0, 0};
                    const int src = (a.slot_mask[sidx] >> fp) & 1;
                    const int nbc = a.slot_nb[sidx * 2 + src];
                    const int blk = a.slot_blk[sidx * 2 + src];
                    const int row = blk * NFP + fp;
                    if (nbc >= 0 && row < a.n_rows) {
                        const real* rw = a.rows + (int64_t)row * NS;
                        const real* qn_src = a.Q + (int64_t)nbc * NS * 5;
                        for (int s = 0; s < NS; ++s) {
                            const real w = rw[s];
#pragma unroll
                            for (int v = 0; v < 5; ++v) dqn[v] += w * qn_src[s * 5 + v];
                        }
                    }
#pragma unroll
                    for (int v = 0; v < 5; ++v) qn[v] = a.qref[v] + (double)dqn[v];
                } else {
                    ghost_state(bc, qo, tn, a.bcp + (a.cell_case ? 40 * a.cell_case[gc] : 0), qn);
                }
                const real* ar = a.slot_adjrow + (sidx * NFP + fp) * 3;
                double A0 = ar[0], A1 = ar[1], A2 = ar[2];
                double mag = rmax(sqrt(A0 * A0 + A1 * A1 + A2 * A2), 1e-300);
                double d[3];
                if ((a.slot_fb[sidx] >> fp) & 1) {
                    d[0] = tn[0]; d[1] = tn[1]; d[2] = tn[2];
                } else {
                    d[0] = A0 / mag * side; d[1] = A1 / mag * side; d[2] = A2 / mag * side;
                }
                double fc[5];
                ausm_up(qo, qn, d, fc);
                double F[3][5];
                euler_flux(qo, F);
                const double sc = mag * side;
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    s_J[((c * NSLOT + slot) * NFP + fp) * 5 + v] =
                        (real)(fc[v] * sc - (A0 * F[0][v] + A1 * F[1][v] + A2 * F[2][v]));
            }
        }
        __syncthreads();

        // 4. volume divergence (folded restriction) + interface correction
        for (int o = t; o < CPB * NS * 5; o += kThreads) {
            const int c = o / (NS * 5), rem = o % (NS * 5), s = rem / 5, v = rem % 5;
            real val = real(0);
            if (c < ncell) {
                real acc = real(0);
                const real* opr = s_op + s * 3 * NF;
                const real* ft = s_Ft + (c * 3 * NF) * 5 + v;
                for (int k = 0; k < 3 * NF; ++k) acc += opr[k] * ft[k * 5];
                const real* gc = a.gcl + ((first + c) * NS + s) * 3;
                acc += gc[0] * (real)a.fref[v] + gc[1] * (real)a.fref[5 + v] + gc[2] * (real)a.fref[10 + v];
                real corr = real(0);
#pragma unroll
                for (int slot = 0; slot < NSLOT; ++slot) {
                    const int code = Cell<CT>::code(slot);
                    const int axis = code >> 1;
                    const real g = s_g[((code & 1) ? N1 : 0) + s_co[axis * NS + s]];
                    corr += g * s_J[((c * NSLOT + slot) * NFP + s_fp[axis * NS + s]) * 5 + v];
                }
                const real dj = a.det[(first + c) * NS + s];
                val = -acc / dj - corr / dj;
            }
            s_R[o] = val;
        }
        __syncthreads();

        // 5. outlier suppression per (cell, variable)
        if (a.suppress && t < CPB * 5) {
            const int c = t / 5, v = t % 5;
            if (c < ncell) {
                real med = median_abs<real, NS>(s_R + c * NS * 5 + v, 5);
                const real mu = a.umean[(first + c) * 5 + v];
                real ref = nan_max(nan_max(med, real(1e-9) * mu), Tiny<real>::v);
                real thr = real(1e4) * ref;
                for (int s = 0; s < NS; ++s) {
                    real x = s_R[(c * NS + s) * 5 + v];
                    if ((x < real(0) ? -x : x) > thr) s_R[(c * NS + s) * 5 + v] = real(0);
                }
            }
        }
        __syncthreads();

        for (int i = t; i < ncell * NS * 5; i += kThreads) a.R[first * NS * 5 + i] = s_R[i];
        __syncthreads();
    }
}

int grid_for(int64_t tiles, int cap = 0) {
    const int64_t c = cap > 0 ? cap : 4096;
    int64_t g = tiles < c ? tiles : c;
    return (int)(g > 0 ? g : 1);
}

template <int P, typename real>
int launch_inviscid(const AfrMeshDesc& d, const real* Q, const real* umean, real* R, int suppress, hipStream_t st) {
    using O = Ord<P>;
    double fref[15];
    {
        double F[3][5];
        euler_flux(d.ref_state, F);
        for (int i = 0; i < 3; ++i)
            for (int v = 0; v < 5; ++v) fref[i * 5 + v] = F[i][v];
    }
    if constexpr (P >= 3) {
        if (d.n_prism > 0) return fail("order 3 is supported on tet meshes only");
    } else if (d.n_prism > 0) {
        // prisms: Kronecker-factored wave-per-cell kernel
        afr_prism::PrismArgs<real> p;
        p.n_cells = d.n_prism;
        p.schedule = d.prism_schedule;
        p.grid_cap = d.grid_cap[AFR_K_PRISM_INVISCID];
        p.variant = d.variant[AFR_K_PRISM_INVISCID];
        p.dQ = Q;
        p.umean = umean;
        p.R = R;
        p.det = (const real*)d.det;
        p.adjf = (const real*)d.adj_fine;
        p.gcl = (const real*)d.gcl;
        p.slot_nb = d.slot_nb;
        p.slot_blk = d.slot_blk;
        p.slot_mask = d.slot_mask;
        p.slot_bc = d.slot_bc;
        p.slot_fb = d.slot_fallback;
        p.slot_tn = (const real*)d.slot_tn;
        p.slot_adjrow = (const real*)d.slot_adjrow;
        p.rows_nnz = d.rows_nnz;
        p.rows_idx = d.rows_idx;
        p.rows_w = (const real*)d.rows_w;
        p.knz = d.rows_knz;
        p.n_rows = d.n_rows;
        p.tab = (const real*)d.prism_tab;
        p.gL = (const real*)d.gL;
        p.gR = (const real*)d.gR;
        p.fp_of_sp = d.fp_of_sp;
        p.coord_of_sp = d.coord_of_sp;
        p.bcp = d.bc_params;
        p.cell_case = d.cell_case;
        for (int v = 0; v < 5; ++v) p.qref[v] = d.ref_state[v];
        for (int k = 0; k < 15; ++k) p.fref[k] = fref[k];
        p.suppress = suppress;
        if (afr_prism::launch_prism_inviscid<P, real>(p, st)) return fail("k_prism_inviscid launch failed");
    }
    // tets: tile-per-workgroup kernel with dense operators
    for (int ct = 1; ct < 2; ++ct) {
        const int64_t begin = ct == 0 ? 0 : d.n_prism;
        const int64_t n = ct == 0 ? d.n_prism : d.n_cells - d.n_prism;
        if (n <= 0) continue;
        Args<real> a;
        a.cell_begin = begin;
        a.n_cells_type = n;
        a.Q = Q;
        a.umean = umean;
        a.R = R;
        a.det = (const real*)d.det;
        a.adjf = (const real*)d.adj_fine;
        a.slot_nb = d.slot_nb;
        a.slot_blk = d.slot_blk;
        a.slot_mask = d.slot_mask;
        a.slot_bc = d.slot_bc;
        a.slot_fb = d.slot_fallback;
        a.slot_tn = (const real*)d.slot_tn;
        a.slot_adjrow = (const real*)d.slot_adjrow;
        a.rows = (const real*)d.rows;
        a.c2f = (const real*)d.c2f[ct];
        a.op = (const real*)d.op[ct];
        a.E = (const real*)d.E[ct];
        a.gL = (const real*)d.gL;
        a.gR = (const real*)d.gR;
        a.fp_of_sp = d.fp_of_sp;
        a.coord_of_sp = d.coord_of_sp;
        a.gcl = (const real*)d.gcl;
        a.bcp = d.bc_params;
        a.cell_case = d.cell_case;
        for (int v = 0; v < 5; ++v) a.qref[v] = d.ref_state[v];
        for (int k = 0; k < 15; ++k) a.fref[k] = fref[k];
        a.suppress = suppress;
        a.n_rows = d.n_rows;
        const int grid = grid_for((n + O::CPB - 1) / O::CPB, d.grid_cap[AFR_K_TET_INVISCID]);
        hipLaunchKernelGGL((k_inviscid<P, 1, real>), dim3(grid), dim3(kThreads), 0, st, a);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

template <typename real>
int launch_primitives(const void* U, void* Q, int64_t n, const double* ref, hipStream_t st) {
    int64_t blocks = (n + kThreads - 1) / kThreads;
    int grid = (int)(blocks < 65536 ? (blocks > 0 ? blocks : 1) : 65536);
    Ref5 r;
    for (int v = 0; v < 5; ++v) r.q[v] = ref ? ref[v] : 0.0;
    hipLaunchKernelGGL((k_primitives<real>), dim3(grid), dim3(kThreads), 0, st, (const real*)U, (real*)Q, n, r);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int afr_version(void) { return AFR_ABI_VERSION; }

int afr_last_error(char* buf, size_t n) {
    if (!buf || n == 0) return -1;
    std::strncpy(buf, g_err.c_str(), n - 1);
    buf[n - 1] = 0;
    return 0;
}

int afr_mesh_create(const AfrMeshDesc* desc, AfrMesh** out) {
    if (!desc || !out) return fail("afr_mesh_create: null argument");
    if (desc->abi_version != AFR_ABI_VERSION) return fail("afr_mesh_create: ABI version mismatch");
    if (desc->order < 1 || desc->order > 3) return fail("afr_mesh_create: order must be 1, 2 or 3");
    if (desc->order == 3 && desc->n_prism > 0) return fail("afr_mesh_create: order 3 is supported on tet meshes only");
    // the collapsed tet operators reach 6e7 at P3: FP32 rounding of the state alone is amplified
    // into O(1e-2) relative residual errors, so P3 runs in FP64 storage and arithmetic
    if (desc->order == 3 && desc->dtype != AFR_F64) return fail("afr_mesh_create: order 3 requires FP64 (dtype 1)");
    if (desc->dtype != AFR_F32 && desc->dtype != AFR_F64) return fail("afr_mesh_create: bad dtype");
    if (desc->n_cells <= 0 || desc->n_prism < 0 || desc->n_prism > desc->n_cells)
        return fail("afr_mesh_create: bad cell counts");
    if (!desc->det || !desc->adj_fine || !desc->slot_nb || !desc->slot_blk || !desc->slot_mask || !desc->slot_bc ||
        !desc->slot_fallback || !desc->slot_tn || !desc->slot_adjrow || !desc->rows || !desc->gL || !desc->gR ||
        !desc->fp_of_sp || !desc->coord_of_sp || !desc->bc_params || !desc->work_q || !desc->gcl ||
        !desc->work_umean || !desc->prism_tab || !desc->rows_nnz || !desc->rows_idx || !desc->rows_w ||
        (desc->n_prism > 0 && !desc->prism_schedule))
        return fail("afr_mesh_create: missing device array");
    {
        const int n1 = desc->order + 1;
        if (desc->rows_knz != n1 * n1 && desc->rows_knz != n1 * n1 * n1)
            return fail("afr_mesh_create: rows_knz must be (order+1)^2 or (order+1)^3 (zero-padded rows)");
    }
    std::vector<int64_t> po, to, wo;
    if (desc->cell_case) {
        const int nc = desc->n_cases;
        if (nc < 1) return fail("afr_mesh_create: cell_case set with n_cases < 1");
        if (!desc->case_prism_off || !desc->case_tet_off) return fail("afr_mesh_create: batched mesh without case offsets");
        po.assign(desc->case_prism_off, desc->case_prism_off + nc + 1);
        to.assign(desc->case_tet_off, desc->case_tet_off + nc + 1);
        if (po[0] != 0 || to[0] != 0 || po[nc] != desc->n_prism || to[nc] != desc->n_cells - desc->n_prism)
            return fail("afr_mesh_create: case offsets must cover the prisms and the tets");
        if (desc->case_wall_off) {
            wo.assign(desc->case_wall_off, desc->case_wall_off + nc + 1);
            if (wo[0] != 0 || wo[nc] != desc->n_wall_rec)
                return fail("afr_mesh_create: case wall offsets must cover the wall records");
        }
        for (int i = 0; i < nc; ++i)
            if (po[i + 1] < po[i] || to[i + 1] < to[i] || (!wo.empty() && wo[i + 1] < wo[i]))
                return fail("afr_mesh_create: case offsets not ascending");
    }
    afr::StateOff so{};
    if (desc->state_ref) {
        if (desc->cell_case) {
            so.table = desc->state_ref;
            so.cell_case = desc->cell_case;
            so.mode = 2;
            // a wave may share one row when cell_case ascends over the prisms and over the tets
            // (the concatenated layout: every case's prisms, then every case's tets)
            std::vector<int> cc((size_t)desc->n_cells);
            if (!cc.empty() && hipMemcpy(cc.data(), desc->cell_case, cc.size() * sizeof(int),
                                         hipMemcpyDeviceToHost) != hipSuccess)
                return fail("afr_mesh_create: cell_case is not a readable device array of n_cells ints");
            bool asc = true;
            for (int64_t c = 1; c < desc->n_cells && asc; ++c)
                if (c != desc->n_prism && cc[c] < cc[c - 1]) asc = false;
            so.sorted = asc ? 1 : 0;
            so.split = desc->n_prism;
        } else {
            if (hipMemcpy(so.one, desc->state_ref, sizeof(so.one), hipMemcpyDeviceToHost) != hipSuccess)
                return fail("afr_mesh_create: state_ref is not a readable device array of 5 doubles");
            so.mode = 1;
        }
    }
    AfrMesh* m = new AfrMesh;
    m->d = *desc;
    m->so = so;
    m->case_prism_off = std::move(po);
    m->case_tet_off = std::move(to);
    m->case_wall_off = std::move(wo);
    // the host arrays are the caller's: only the copies are kept
    m->d.case_prism_off = m->d.case_tet_off = m->d.case_wall_off = nullptr;
    for (int k = 0; k < 8; ++k) m->d.grid_cap[k] = m->d.variant[k] = 0;
    *out = m;
    return 0;
}

int afr_mesh_set_grid(AfrMesh* mesh, int32_t kernel, int32_t grid_cap) {
    if (!mesh) return fail("afr_mesh_set_grid: null mesh");
    if (kernel < 0 || kernel > AFR_K_SST_UPDATE) return fail("afr_mesh_set_grid: unknown kernel id");
    if (grid_cap != 0 && (grid_cap < 8 || grid_cap > 65536))
        return fail("afr_mesh_set_grid: grid_cap must be 0 or in [8, 65536]");
    mesh->d.grid_cap[kernel] = grid_cap / 8 * 8;
    return 0;
}

int afr_mesh_get_grid(const AfrMesh* mesh, int32_t kernel) {
    if (!mesh || kernel < 0 || kernel > AFR_K_SST_UPDATE) return fail("afr_mesh_get_grid: bad argument");
    return mesh->d.grid_cap[kernel];
}

int afr_mesh_set_variant(AfrMesh* mesh, int32_t kernel, int32_t variant) {
    if (!mesh) return fail("afr_mesh_set_variant: null mesh");
    if (kernel < 0 || kernel > AFR_K_SST_UPDATE) return fail("afr_mesh_set_variant: unknown kernel id");
    const int n_var = (kernel == AFR_K_PRISM_INVISCID || kernel == AFR_K_SST_CELL) ? 2 : 1;
    if (variant < 0 || variant >= n_var) return fail("afr_mesh_set_variant: no such variant of this kernel");
    mesh->d.variant[kernel] = variant;
    return 0;
}

int afr_mesh_get_variant(const AfrMesh* mesh, int32_t kernel) {
    if (!mesh || kernel < 0 || kernel > AFR_K_SST_UPDATE) return fail("afr_mesh_get_variant: bad argument");
    return mesh->d.variant[kernel];
}

int afr_mesh_destroy(AfrMesh* mesh) {
    delete mesh;
    return 0;
}

struct AfrGraph {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

int afr_graph_begin(void* stream) {
    if (!stream) return fail("afr_graph_begin: capture needs a created stream, not the null stream");
    HIP_TRY(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return 0;
}

int afr_graph_end(void* stream, AfrGraph** out) {
    if (!stream || !out) return fail("afr_graph_end: null stream or output");
    *out = nullptr;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t e = nullptr;
    const hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    if (r != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail((std::string("afr_graph_end: hipGraphInstantiate: ") + hipGetErrorString(r)).c_str());
    }
    *out = new AfrGraph{g, e};
    return 0;
}

int afr_graph_launch(AfrGraph* graph, void* stream) {
    if (!graph) return fail("afr_graph_launch: null graph");
    HIP_TRY(hipGraphLaunch(graph->exec, (hipStream_t)stream));
    return 0;
}

int afr_graph_destroy(AfrGraph* graph) {
    if (!graph) return 0;
    const hipError_t r1 = hipGraphExecDestroy(graph->exec);
    const hipError_t r2 = hipGraphDestroy(graph->graph);
    delete graph;
    if (r1 != hipSuccess || r2 != hipSuccess) return fail("afr_graph_destroy: hipGraph(Exec)Destroy failed");
    return 0;
}

int afr_primitives(const void* U, void* Q, int64_t n_points, int32_t dtype, const double* ref, void* stream) {
    if (!U || !Q || n_points < 0) return fail("afr_primitives: bad argument");
    hipStream_t st = (hipStream_t)stream;
    return dtype == AFR_F64 ? launch_primitives<double>(U, Q, n_points, ref, st)
                            : launch_primitives<float>(U, Q, n_points, ref, st);
}

int afr_prepare_state(const AfrMesh* mesh, const void* U, void* dQ, void* umean, void* stream) {
    if (!mesh || !U || !dQ || !umean) return fail("afr_prepare_state: null argument");
    const AfrMeshDesc& d = mesh->d;
    hipStream_t st = (hipStream_t)stream;
    const int ns = (d.order + 1) * (d.order + 1) * (d.order + 1);
    int rc;
    if (d.dtype == AFR_F32) {
        auto* u = (const float*)U;
        auto* q = (float*)dQ;
        auto* m = (float*)umean;
        rc = ns == 8    ? afr_prism::launch_prim_umean<8, float>(u, q, m, d.n_cells, d.ref_state, mesh->so, st)
             : ns == 27 ? afr_prism::launch_prim_umean<27, float>(u, q, m, d.n_cells, d.ref_state, mesh->so, st)
                        : afr_prism::launch_prim_umean<64, float>(u, q, m, d.n_cells, d.ref_state, mesh->so, st);
    } else {
        auto* u = (const double*)U;
        auto* q = (double*)dQ;
        auto* m = (double*)umean;
        rc = ns == 8    ? afr_prism::launch_prim_umean<8, double>(u, q, m, d.n_cells, d.ref_state, mesh->so, st)
             : ns == 27 ? afr_prism::launch_prim_umean<27, double>(u, q, m, d.n_cells, d.ref_state, mesh->so, st)
                        : afr_prism::launch_prim_umean<64, double>(u, q, m, d.n_cells, d.ref_state, mesh->so, st);
    }
    return rc ? fail("k_prim_umean launch failed") : 0;
}

int afr_inviscid_from_primitives(const AfrMesh* mesh, const void* dQ, const void* umean, void* R,
                                 int32_t suppress, void* stream) {
    if (!mesh || !dQ || !umean || !R) return fail("afr_inviscid_from_primitives: null argument");
    const AfrMeshDesc& d = mesh->d;
    hipStream_t st = (hipStream_t)stream;
    if (d.dtype == AFR_F32) {
        auto* q = (const float*)dQ;
        auto* m = (const float*)umean;
        return d.order == 1   ? launch_inviscid<1, float>(d, q, m, (float*)R, suppress, st)
               : d.order == 2 ? launch_inviscid<2, float>(d, q, m, (float*)R, suppress, st)
                              : launch_inviscid<3, float>(d, q, m, (float*)R, suppress, st);
    }
    auto* q = (const double*)dQ;
    auto* m = (const double*)umean;
    return d.order == 1   ? launch_inviscid<1, double>(d, q, m, (double*)R, suppress, st)
           : d.order == 2 ? launch_inviscid<2, double>(d, q, m, (double*)R, suppress, st)
                          : launch_inviscid<3, double>(d, q, m, (double*)R, suppress, st);
}

int afr_inviscid_residual(const AfrMesh* mesh, const void* U, void* R, int32_t suppress, void* stream) {
    if (!mesh || !U || !R) return fail("afr_inviscid_residual: null argument");
    const AfrMeshDesc& d = mesh->d;
    if (afr_prepare_state(mesh, U, d.work_q, d.work_umean, stream)) return -1;
    return afr_inviscid_from_primitives(mesh, d.work_q, d.work_umean, R, suppress, stream);
}

}  // extern "C"

namespace {

template <typename real>
int run_viscous(const AfrMeshDesc& d, const StateOff& so, const void* U, const void* mu_t, void* R, double mu,
                double Pr, double Prt, int suppress, int accumulate, int phase, hipStream_t st) {
    // phase 0: volume kernel only; 1: interface kernel only
    for (int ct = 0; ct < 2; ++ct) {
        afr_visc::ViscArgs<real> a;
        a.cell_begin = ct == 0 ? 0 : d.n_prism;
        a.n_cells = ct == 0 ? d.n_prism : d.n_cells - d.n_prism;
        if (a.n_cells <= 0) continue;
        a.U = (const real*)U;
        a.so = so;
        a.mu_t = (const real*)mu_t;
        a.dQ = (real*)d.work_q;
        a.umean = (real*)d.work_umean;
        a.grad = (real*)d.work_grad;
        a.rvol = (real*)d.work_rvol;
        a.h_ip = (real*)d.work_h;
        a.R = (real*)R;
        a.det = (const real*)d.det;
        a.adj = (const real*)d.adj;
        a.vpack = (real*)d.work_vpack;
        // the interface kernel reads neighbour data: locality order; the volume kernel only
        // its own cell: storage order (contiguous reads)
        a.schedule = (ct == 0 && phase != 0) ? d.prism_schedule : nullptr;
        a.grid_cap = d.grid_cap[phase == 0 ? AFR_K_VISC_VOLUME : AFR_K_VISC_IFACE];
        a.Dsp_w = (const real*)d.dsp_w[ct];
        a.Dsp_i = d.dsp_i[ct];
        a.Esp_w = (const real*)d.esp_w[ct];
        a.Esp_i = d.esp_i[ct];
        a.slot_nb = d.slot_nb;
        a.slot_blk = d.slot_blk;
        a.slot_mask = d.slot_mask;
        a.slot_bc = d.slot_bc;
        a.slot_tn = (const real*)d.slot_tn;
        a.slot_adjrow = (const real*)d.slot_adjrow;
        a.rows_idx = d.rows_idx;
        a.rows_w = (const real*)d.rows_w;
        a.knz = d.rows_knz;
        a.n_rows = d.n_rows;
        a.gL = (const real*)d.gL;
        a.gR = (const real*)d.gR;
        a.fp_of_sp = d.fp_of_sp;
        a.coord_of_sp = d.coord_of_sp;
        a.bcp = d.bc_params;
        a.cell_case = d.cell_case;
        for (int v = 0; v < 5; ++v) a.qref[v] = d.ref_state[v];
        a.mu = mu;
        a.Pr = Pr;
        a.Prt = Prt;
        a.suppress = suppress;
        a.accumulate = accumulate;
        int rc;
        if (d.order == 3)
            rc = ct == 0 ? -1 : afr_visc::launch_viscous_phase<3, 1, real>(a, phase, st);
        else if (d.order == 1)
            rc = ct == 0 ? afr_visc::launch_viscous_phase<1, 0, real>(a, phase, st)
                         : afr_visc::launch_viscous_phase<1, 1, real>(a, phase, st);
        else
            rc = ct == 0 ? afr_visc::launch_viscous_phase<2, 0, real>(a, phase, st)
                         : afr_visc::launch_viscous_phase<2, 1, real>(a, phase, st);
        if (rc) return fail("viscous kernel launch failed");
    }
    return 0;
}

int viscous_phase(const AfrMesh* mesh, const void* U, const void* mu_t, void* R, double mu, double Pr, double Prt,
                  int suppress, int accumulate, int phase, hipStream_t st) {
    const AfrMeshDesc& d = mesh->d;
    return d.dtype == AFR_F32
               ? run_viscous<float>(d, mesh->so, U, mu_t, R, mu, Pr, Prt, suppress, accumulate, phase, st)
               : run_viscous<double>(d, mesh->so, U, mu_t, R, mu, Pr, Prt, suppress, accumulate, phase, st);
}

bool viscous_ready(const AfrMeshDesc& d) {
    return d.adj && d.work_grad && d.work_rvol && d.work_h && d.work_vpack && d.dsp_w[0] && d.dsp_w[1] && d.dsp_i[0] &&
           d.dsp_i[1] && d.esp_w[0] && d.esp_w[1] && d.esp_i[0] && d.esp_i[1];
}

}  // namespace

extern "C" {

int afr_viscous_residual(const AfrMesh* mesh, const void* U, const void* mu_t, void* R, double mu, double Pr,
                         double Pr_t, int32_t suppress, void* stream) {
    if (!mesh || !U || !R) return fail("afr_viscous_residual: null argument");
    const AfrMeshDesc& d = mesh->d;
    if (!viscous_ready(d)) return fail("afr_viscous_residual: mesh has no viscous work buffers");
    hipStream_t st = (hipStream_t)stream;
    if (viscous_phase(mesh, U, mu_t, R, mu, Pr, Pr_t, suppress, 0, 0, st)) return -1;
    return viscous_phase(mesh, U, mu_t, R, mu, Pr, Pr_t, suppress, 0, 1, st);
}

int afr_viscous_volume(const AfrMesh* mesh, const void* U, const void* mu_t, double mu, double Pr, double Pr_t,
                       void* stream) {
    if (!mesh || !U) return fail("afr_viscous_volume: null argument");
    const AfrMeshDesc& d = mesh->d;
    if (!viscous_ready(d)) return fail("afr_viscous_volume: mesh has no viscous work buffers");
    return viscous_phase(mesh, U, mu_t, nullptr, mu, Pr, Pr_t, 0, 0, 0, (hipStream_t)stream);
}

int afr_viscous_interface(const AfrMesh* mesh, const void* U, const void* mu_t, void* R, double mu, double Pr,
                          double Pr_t, int32_t suppress, int32_t accumulate, void* stream) {
    if (!mesh || !U || !R) return fail("afr_viscous_interface: null argument");
    const AfrMeshDesc& d = mesh->d;
    if (!viscous_ready(d)) return fail("afr_viscous_interface: mesh has no viscous work buffers");
    return viscous_phase(mesh, U, mu_t, R, mu, Pr, Pr_t, suppress, accumulate, 1, (hipStream_t)stream);
}

int afr_residual(const AfrMesh* mesh, const void* U, const void* mu_t, void* R, double mu, double Pr, double Pr_t,
                 int32_t suppress, void* stream) {
    if (!mesh || !U || !R) return fail("afr_residual: null argument");
    const AfrMeshDesc& d = mesh->d;
    if (!viscous_ready(d)) return fail("afr_residual: mesh has no viscous work buffers");
    hipStream_t st = (hipStream_t)stream;
    // volume kernel: prepared state (dQ, umean), gradients, viscous volume term
    if (viscous_phase(mesh, U, mu_t, R, mu, Pr, Pr_t, suppress, 0, 0, st)) return -1;
    if (afr_inviscid_from_primitives(mesh, d.work_q, d.work_umean, R, suppress, stream)) return -1;
    return viscous_phase(mesh, U, mu_t, R, mu, Pr, Pr_t, suppress, 1, 1, st);
}

namespace {

bool step_ready(const AfrMeshDesc& d) {
    return d.metric_flux_scale && d.volume && d.work_r && d.work_s1 && d.work_s2 && d.work_dt && d.work_part &&
           d.filt[0] && d.filt[1];
}

int ns_of(const AfrMeshDesc& d) { return (d.order + 1) * (d.order + 1) * (d.order + 1); }

}  // namespace

static int local_time_step(const AfrMesh* mesh, const void* U, const void* mu_t, void* dt, double mu, int nut_mode,
                           void* stream) {
    if (!mesh || !U || !dt) return fail("afr_local_time_step: null argument");
    const AfrMeshDesc& d = mesh->d;
    if (!d.metric_flux_scale || !d.volume) return fail("afr_local_time_step: mesh has no metric_flux_scale/volume");
    hipStream_t st = (hipStream_t)stream;
    const int rc = d.dtype == AFR_F32
                       ? afr_step::launch_dt_local<float>((const float*)U, mesh->so, (const float*)mu_t, (const float*)d.det,
                                                          (const float*)d.metric_flux_scale, (const float*)d.volume,
                                                          (float*)dt, d.n_cells, ns_of(d), mu, d.order, nut_mode, st)
                       : afr_step::launch_dt_local<double>((const double*)U, mesh->so, (const double*)mu_t, (const double*)d.det,
                                                           (const double*)d.metric_flux_scale,
                                                           (const double*)d.volume, (double*)dt, d.n_cells, ns_of(d),
                                                           mu, d.order, nut_mode, st);
    return rc ? fail("k_dt_local launch failed") : 0;
}

int afr_local_time_step(const AfrMesh* mesh, const void* U, const void* mu_t, void* dt, double mu, void* stream) {
    return local_time_step(mesh, U, mu_t, dt, mu, 0, stream);
}

int afr_local_time_step_nut(const AfrMesh* mesh, const void* U, const void* nu_t, void* dt, double mu, void* stream) {
    return local_time_step(mesh, U, nu_t, dt, mu, 1, stream);
}

int afr_rk_stage(const AfrMesh* mesh, const void* U0, const void* Ua, const void* R, const void* dt, void* U_out,
                 double c0, double c1, double c2, void* stream) {
    if (!mesh || !U0 || !R || !dt || !U_out || (c1 != 0.0 && !Ua)) return fail("afr_rk_stage: null argument");
    const AfrMeshDesc& d = mesh->d;
    if (!d.filt[0] || !d.filt[1]) return fail("afr_rk_stage: mesh has no filter tables");
    hipStream_t st = (hipStream_t)stream;
    const void* ua = Ua ? Ua : U0;
    const int rc = d.dtype == AFR_F32
                       ? afr_step::launch_rk_stage<float>(ns_of(d), (const float*)U0, (const float*)ua, (const float*)R,
                                                          (const float*)dt, (float*)U_out, mesh->so, (const float*)d.filt[0],
                                                          (const float*)d.filt[1], d.n_cells, d.n_prism, c0, c1, c2,
                                                          1.0, st)
                       : afr_step::launch_rk_stage<double>(ns_of(d), (const double*)U0, (const double*)ua,
                                                           (const double*)R, (const double*)dt, (double*)U_out,
                                                           mesh->so, (const double*)d.filt[0], (const double*)d.filt[1],
                                                           d.n_cells, d.n_prism, c0, c1, c2, 1.0, st);
    return rc ? fail("k_rk_stage launch failed") : 0;
}

int afr_residual_norm_f64(const AfrMesh* mesh, const void* R, int32_t n_vars_total, double* out, void* stream) {
    if (!mesh || !R || !out) return fail("afr_residual_norm_f64: null argument");
    const AfrMeshDesc& d = mesh->d;
    if (!d.work_part) return fail("afr_residual_norm_f64: mesh has no work_part buffer");
    if (n_vars_total < 5) return fail("afr_residual_norm_f64: n_vars_total < 5");
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = d.n_cells * ns_of(d) * 5;
    const double n_total = (double)d.n_cells * ns_of(d) * n_vars_total;
    const int rc = d.dtype == AFR_F32
                       ? afr_step::launch_norm<float>((const float*)R, n, n_total, d.work_part, afr_step::kNormParts,
                                                      out, st)
                       : afr_step::launch_norm<double>((const double*)R, n, n_total, d.work_part,
                                                       afr_step::kNormParts, out, st);
    return rc ? fail("residual norm launch failed") : 0;
}

}  // extern "C"

// Partial sums of afr_residual_norm_f64_noalloc: a file-scope device array, one copy per GPU, held
// until the process ends.  Host code never passes `g_norm_scratch` itself (in host code the name
// is a host-side stand-in, not a device address): the device address comes from
// hipGetSymbolAddress, once per device, with that device current.
static __device__ __align__(256) double g_norm_scratch[afr_step::kNormParts];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// device address of this GPU's g_norm_scratch (nullptr on failure, with the error set)
double* norm_scratch_of_current_device() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        fail("afr_residual_norm_f64_noalloc: no current HIP device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(g_norm_scratch)) != hipSuccess || !ptr) {
            fail("afr_residual_norm_f64_noalloc: hipGetSymbolAddress(g_norm_scratch) failed");
            return nullptr;
        }
        g_scratch_ptr[dev] = (double*)ptr;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int afr_residual_norm_f64_noalloc(const AfrMesh* mesh, const void* R, int32_t n_vars_total, double* out,
                                  void* stream) {
    if (!mesh || !R || !out) return fail("afr_residual_norm_f64_noalloc: null argument");
    const AfrMeshDesc& d = mesh->d;
    if (n_vars_total < 5) return fail("afr_residual_norm_f64_noalloc: n_vars_total < 5");
    double* part = norm_scratch_of_current_device();
    if (!part) return -1;
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = d.n_cells * ns_of(d) * 5;
    const double n_total = (double)d.n_cells * ns_of(d) * n_vars_total;
    // the partition is the one of afr_residual_norm_f64: at most kNormParts partial sums, whatever n
    const int rc = d.dtype == AFR_F32
                       ? afr_step::launch_norm<float>((const float*)R, n, n_total, part, afr_step::kNormParts, out,
                                                      st)
                       : afr_step::launch_norm<double>((const double*)R, n, n_total, part, afr_step::kNormParts,
                                                       out, st);
    return rc ? fail("residual norm launch failed") : 0;
}

int afr_residual_norm_cases(const AfrMesh* mesh, const void* R, int32_t n_vars_total, double* out, void* stream) {
    if (!mesh || !R || !out) return fail("afr_residual_norm_cases: null argument");
    const AfrMeshDesc& d = mesh->d;
    if (!d.work_part) return fail("afr_residual_norm_cases: mesh has no work_part buffer");
    if (n_vars_total < 5) return fail("afr_residual_norm_cases: n_vars_total < 5");
    if (mesh->case_prism_off.empty()) return afr_residual_norm_f64(mesh, R, n_vars_total, out, stream);
    hipStream_t st = (hipStream_t)stream;
    const int64_t cv = (int64_t)ns_of(d) * 5;
    const auto& po = mesh->case_prism_off;
    const auto& to = mesh->case_tet_off;
    for (int i = 0; i < d.n_cases; ++i) {
        // the work_part buffer is reused case after case (stream order)
        const int64_t a0 = po[i] * cv, n0 = (po[i + 1] - po[i]) * cv;
        const int64_t a1 = (d.n_prism + to[i]) * cv, n1 = (to[i + 1] - to[i]) * cv;
        const double n_total = (double)((n0 + n1) / 5) * n_vars_total;
        const int rc = d.dtype == AFR_F32
                           ? afr_step::launch_norm2<float>((const float*)R, a0, n0, a1, n1, n_total, d.work_part,
                                                           afr_step::kNormParts, out + i, st)
                           : afr_step::launch_norm2<double>((const double*)R, a0, n0, a1, n1, n_total, d.work_part,
                                                            afr_step::kNormParts, out + i, st);
        if (rc) return fail("residual norm launch failed");
    }
    return 0;
}

static int rk3_stages(const AfrMesh* mesh, void* U, const void* mu_t, double mu, double Pr, double Pr_t,
                      int32_t n_vars_total, double* norm, void* stream) {
    const AfrMeshDesc& d = mesh->d;
    const bool visc = mu != 0.0 || mu_t != nullptr;
    auto rhs = [&](const void* X, void* R) -> int {
        return visc ? afr_residual(mesh, X, mu_t, R, mu, Pr, Pr_t, 1, stream)
                    : afr_inviscid_residual(mesh, X, R, 1, stream);
    };
    if (rhs(U, d.work_r)) return -1;
    if (afr_residual_norm_cases(mesh, d.work_r, n_vars_total, norm, stream)) return -1;
    // Shu-Osher SSP-RK3: U1 = U0 + dt L0; U2 = 3/4 U0 + 1/4 U1 + 1/4 dt L1; U3 = 1/3 U0 + 2/3 U2 + 2/3 dt L2
    if (afr_rk_stage(mesh, U, nullptr, d.work_r, d.work_dt, d.work_s1, 1.0, 0.0, 1.0, stream)) return -1;
    if (rhs(d.work_s1, d.work_r)) return -1;
    if (afr_rk_stage(mesh, U, d.work_s1, d.work_r, d.work_dt, d.work_s2, 0.75, 0.25, 0.25, stream)) return -1;
    if (rhs(d.work_s2, d.work_r)) return -1;
    return afr_rk_stage(mesh, U, d.work_s2, d.work_r, d.work_dt, U, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, stream);
}

int afr_ssp_rk3_step(const AfrMesh* mesh, void* U, const void* mu_t, double mu, double Pr, double Pr_t,
                     int32_t n_vars_total, double* norm, void* stream) {
    if (!mesh || !U || !norm) return fail("afr_ssp_rk3_step: null argument");
    const AfrMeshDesc& d = mesh->d;
    if (!step_ready(d)) return fail("afr_ssp_rk3_step: mesh has no pseudo-step work buffers");
    if (afr_local_time_step(mesh, U, mu_t, d.work_dt, mu, stream)) return -1;
    return rk3_stages(mesh, U, mu_t, mu, Pr, Pr_t, n_vars_total, norm, stream);
}

int afr_ssp_rk3_step_dt(const AfrMesh* mesh, void* U, const void* mu_t, double dt, double mu, double Pr, double Pr_t,
                        int32_t n_vars_total, double* norm, void* stream) {
    if (!mesh || !U || !norm) return fail("afr_ssp_rk3_step_dt: null argument");
    const AfrMeshDesc& d = mesh->d;
    if (!step_ready(d)) return fail("afr_ssp_rk3_step_dt: mesh has no pseudo-step work buffers");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail("afr_ssp_rk3_step_dt: dt must be positive and finite");
    const int64_t n = d.n_cells * ns_of(d);
    const int rc = d.dtype == AFR_F32 ? afr_step::launch_fill<float>((float*)d.work_dt, n, dt, (hipStream_t)stream)
                                      : afr_step::launch_fill<double>((double*)d.work_dt, n, dt, (hipStream_t)stream);
    if (rc) return fail("k_fill launch failed");
    return rk3_stages(mesh, U, mu_t, mu, Pr, Pr_t, n_vars_total, norm, stream);
}

int afr_sst_rk3_step(const AfrMesh* mesh, void* U, void* k, void* omega, const void* d_wall, void* nu_t, void* mu_t,
                     double mu, double Pr, double Pr_t, int32_t n_vars_total, double* norm, void* stream) {
    if (!mesh || !U || !k || !omega || !d_wall || !nu_t || !mu_t || !norm) return fail("afr_sst_rk3_step: null argument");
    const AfrMeshDesc& d = mesh->d;
    if (!step_ready(d)) return fail("afr_sst_rk3_step: mesh has no pseudo-step work buffers");
    if (afr_local_time_step_nut(mesh, U, nu_t, d.work_dt, mu, stream)) return -1;
    if (afr_sst_step(mesh, U, k, omega, d_wall, d.work_dt, nu_t, mu_t, nullptr, mu, stream)) return -1;
    return rk3_stages(mesh, U, mu_t, mu, Pr, Pr_t, n_vars_total, norm, stream);
}

int afr_wall_distance(const double* sp_xyz, int64_t n_pts, const double* wall_xyz, int32_t n_wall, void* d_out,
                      int32_t dtype, void* stream) {
    if (!sp_xyz || !d_out || n_pts < 0 || n_wall < 1 || !wall_xyz) return fail("afr_wall_distance: bad argument");
    if (dtype != AFR_F32 && dtype != AFR_F64) return fail("afr_wall_distance: bad dtype");
    return afr_sst::launch_wall_distance(sp_xyz, n_pts, wall_xyz, n_wall, d_out, dtype, (hipStream_t)stream)
               ? fail("k_wall_distance launch failed")
               : 0;
}

}  // extern "C"

namespace {

template <typename real>
int run_sst(const AfrMeshDesc& d, const StateOff& so, const void* U, void* k, void* w, const void* dw,
            const void* dt, void* nut, void* mut, void* rates, double mu, hipStream_t st) {
    if (d.order > 2) return fail("SST is supported at orders 1 and 2");
    afr_sst::SstArgs<real> a;
    a.n_records = d.n_records;
    a.U = (const real*)U;
    a.so = so;
    a.k = (real*)k;
    a.w = (real*)w;
    a.d_wall = (const real*)dw;
    a.dt = (const real*)dt;
    a.nu_t = (real*)nut;
    a.mu_t = (real*)mut;
    a.pack = (real*)d.work_pack;
    a.nbp = (real*)d.work_nbp;
    a.rates = (real*)rates;
    a.vol = (real*)d.work_tvol;
    a.src = (real*)d.work_tsrc;
    a.crec = (real*)d.work_crec;
    a.det = (const real*)d.det;
    a.adj = (const real*)d.adj;
    a.E0 = (const real*)d.E[0];
    a.E1 = (const real*)d.E[1];
    a.slot_nb = d.slot_nb;
    a.slot_blk = d.slot_blk;
    a.slot_mask = d.slot_mask;
    a.rows_idx = d.rows_idx;
    a.rows_w = (const real*)d.rows_w;
    a.knz = d.rows_knz;
    a.n_rows = d.n_rows;
    a.n_prism = d.n_prism;
    a.rec_owner = d.rec_owner;
    a.rec_oslot = d.rec_oslot;
    a.rec_flags = d.rec_flags;
    a.rec_tn = (const real*)d.rec_tn;
    a.inc = d.cell_inc;
    a.gL = (const real*)d.gL;
    a.gR = (const real*)d.gR;
    a.fp_of_sp = d.fp_of_sp;
    a.coord_of_sp = d.coord_of_sp;
    a.mu_molecular = mu;
    a.grid_cap_face = d.grid_cap[AFR_K_SST_FACE];
    const int P = d.order;
    for (int phase = 0; phase < 3; ++phase) {
        if (phase == 1) {
            const int rc = P == 1 ? afr_sst::launch_sst_face<1, real>(a, st) : afr_sst::launch_sst_face<2, real>(a, st);
            if (rc) return fail("k_sst_face launch failed");
            continue;
        }
        for (int ct = 0; ct < 2; ++ct) {
            afr_sst::SstArgs<real> b = a;
            b.cell_begin = ct == 0 ? 0 : d.n_prism;
            b.n_cells = ct == 0 ? d.n_prism : d.n_cells - d.n_prism;
            b.D = (const real*)d.D[ct];
            b.Dsp_w = (const real*)d.dsp_w[ct];
            b.Dsp_i = d.dsp_i[ct];
            b.Esp_w = (const real*)d.esp_w[ct];
            b.Esp_i = d.esp_i[ct];
            b.schedule = ct == 0 ? d.prism_schedule : nullptr;
            b.grid_cap = d.grid_cap[phase == 0 ? AFR_K_SST_CELL : AFR_K_SST_UPDATE];
            b.variant = phase == 0 ? d.variant[AFR_K_SST_CELL] : 0;
            if (b.n_cells <= 0) continue;
            int rc;
            if (phase == 0)
                rc = P == 1 ? (ct == 0 ? afr_sst::launch_sst_cell<1, 0, real>(b, st) : afr_sst::launch_sst_cell<1, 1, real>(b, st))
                            : (ct == 0 ? afr_sst::launch_sst_cell<2, 0, real>(b, st) : afr_sst::launch_sst_cell<2, 1, real>(b, st));
            else
                rc = P == 1 ? (ct == 0 ? afr_sst::launch_sst_update<1, 0, real>(b, st)
                                       : afr_sst::launch_sst_update<1, 1, real>(b, st))
                            : (ct == 0 ? afr_sst::launch_sst_update<2, 0, real>(b, st)
                                       : afr_sst::launch_sst_update<2, 1, real>(b, st));
            if (rc) return fail("SST kernel launch failed");
        }
    }
    return 0;
}

}  // namespace

extern "C" {

int afr_sst_step(const AfrMesh* mesh, const void* U, void* k, void* omega, const void* d_wall, const void* dt,
                 void* nu_t, void* mu_t, void* rates, double mu_molecular, void* stream) {
    if (!mesh || !U || !k || !omega || !d_wall || !dt || !nu_t || !mu_t) return fail("afr_sst_step: null argument");
    const AfrMeshDesc& d = mesh->d;
    if (d.n_records <= 0 || !d.rec_owner || !d.rec_oslot || !d.rec_flags || !d.rec_tn || !d.cell_inc ||
        !d.work_pack || !d.work_nbp || !d.work_tvol || !d.work_tsrc || !d.work_crec || !d.adj)
        return fail("afr_sst_step: mesh has no SST record tables / work buffers");
    hipStream_t st = (hipStream_t)stream;
    return d.dtype == AFR_F32
               ? run_sst<float>(d, mesh->so, U, k, omega, d_wall, dt, nu_t, mu_t, rates, mu_molecular, st)
               : run_sst<double>(d, mesh->so, U, k, omega, d_wall, dt, nu_t, mu_t, rates, mu_molecular, st);
}

// wall_off == nullptr: one sum over every record into out[6]; else per case i the records
// [wall_off[i], wall_off[i+1]) into out[6 i .. 6 i + 5]
static int wall_forces(const AfrMesh* mesh, const void* U, const void* mu_t, double mu, bool per_case, double* out,
                       void* stream) {
    if (mesh->d.order > 3 || (mesh->d.order == 3 && mesh->d.dtype != AFR_F64))
        return fail("afr_forces_f64: wall forces are supported at orders 1 and 2, and at 3 in FP64");
    const int64_t* wall_off = per_case ? mesh->case_wall_off.data() : nullptr;
    const int n_cases = mesh->d.n_cases;
    if (per_case && mesh->case_wall_off.empty())
        return fail("afr_forces_cases_f64: mesh has no per-case wall offsets");
    const AfrMeshDesc& d = mesh->d;
    if (d.n_wall_rec < 0 || (d.n_wall_rec > 0 && (!d.wrec_owner || !d.wrec_oslot || !d.wrec_geom || !d.w1d ||
                                                  !d.work_fpart || !d.adj)))
        return fail("afr_forces_f64: mesh has no wall-record tables");
    hipStream_t st = (hipStream_t)stream;
    auto fill = [&](auto* tag) {
        using real = std::remove_pointer_t<decltype(tag)>;
        afr_forces::ForceArgs<real> a;
        a.n_wall = d.n_wall_rec;
        a.n_prism = d.n_prism;
        a.U = (const real*)U;
        a.so = mesh->so;
        a.mu_t = (const real*)mu_t;
        a.det = (const real*)d.det;
        a.adj = (const real*)d.adj;
        a.D0 = (const real*)d.D[0];
        a.D1 = (const real*)d.D[1];
        a.E0 = (const real*)d.E[0];
        a.E1 = (const real*)d.E[1];
        a.wrec_owner = d.wrec_owner;
        a.wrec_oslot = d.wrec_oslot;
        a.wrec_geom = d.wrec_geom;
        a.w1d = d.w1d;
        a.mu = mu;
        a.part = d.work_fpart;
        a.out = out;
        auto records = [&]() {
            if constexpr (std::is_same<real, double>::value) {
                if (d.order == 3) return afr_forces::launch_wall_forces_records<3, double>(a, st);
            }
            return d.order == 1 ? afr_forces::launch_wall_forces_records<1, real>(a, st)
                                : afr_forces::launch_wall_forces_records<2, real>(a, st);
        };
        if (!wall_off) {
            if (records()) return -1;
            return afr_forces::launch_sum_forces(a.part, a.n_wall, a.out, st);
        }
        if (records()) return -1;
        for (int i = 0; i < n_cases; ++i)
            if (afr_forces::launch_sum_forces(d.work_fpart + wall_off[i] * 6, wall_off[i + 1] - wall_off[i],
                                              out + 6 * i, st))
                return -1;
        return 0;
    };
    const int rc = d.dtype == AFR_F32 ? fill((float*)nullptr) : fill((double*)nullptr);
    return rc ? fail("wall force kernel launch failed") : 0;
}

int afr_forces_f64(const AfrMesh* mesh, const void* U, const void* mu_t, double mu, double* out, void* stream) {
    if (!mesh || !U || !out) return fail("afr_forces_f64: null argument");
    return wall_forces(mesh, U, mu_t, mu, false, out, stream);
}

int afr_forces_cases_f64(const AfrMesh* mesh, const void* U, const void* mu_t, double mu, double* out, void* stream) {
    if (!mesh || !U || !out) return fail("afr_forces_cases_f64: null argument");
    return wall_forces(mesh, U, mu_t, mu, true, out, stream);
}

}  // extern "C"
