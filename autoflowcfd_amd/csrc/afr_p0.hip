// The P0 phase of the order continuation and the interpolation between orders.
//
//   k_p0_residual    the P0 finite-volume inviscid residual of
//                    src/autoflowcfd/core/fr_residual/inviscid_p0.py:17-143 (inviscid_p0_kernel.py:24-105):
//                    every face record carries AUSM+up(Q_owner, Q_other, n) x area, minus on the owner,
//                    plus on the neighbour, each over its cell volume (the target-order volume); a
//                    boundary record's other state is the ghost state its owner cell keeps, that of
//                    the cell's last boundary record (inviscid_p0.py:118-143).  Plus the viscous residual
//                    at one solution point (viscous_p0_kernel.py:26-285): with the gradients of a constant
//                    zero, only the interior-penalty term of the owner-primary boundary records remains,
//                    distributed with g'(0) = -1/2 / +1/2 over det J at the point.  One thread per cell
//                    gathers its records (no scatter, no atomics: a record's flux is evaluated by both of
//                    its cells from the same inputs, hence the same bits).
//   k_p0_stage       one Shu-Osher stage with the positivity projection, one thread per cell (the P0
//                    filter is the identity).
//   k_interp_order   X_new[c][a][v] = sum_b W[a][b] X_old[c][b][v]: the tensor-product linear
//                    interpolation of src/autoflowcfd/core/utils/order_continuation.py:14-57.
//   k_p0_sst_point / k_p0_sst_transport / k_p0_sst_update
//                    the SST turbulence step of the P0 phase (the FR turbulence step with the
//                    one-point operators, src/autoflowcfd/core/fr_solver/turbulence.py:178-330): per
//                    cell the point physics with zero gradients (sst.py:203-330), the transport
//                    gathered over the cell's face records (transport.py:122-520,
//                    transport_kernel.py:103-260: every record, owner-side extrapolation, neighbour
//                    sources or the Neumann / Dirichlet-zero ghost, upwind convection, BR1-type
//                    diffusion, distribution with g' = -1/2, +1/2 over det J; the two corrections are
//                    summed separately and subtracted, as the reference does), then the
//                    point-implicit update with floors (sst.py:335-479).
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/afr.h"
#include "afr_physics.hpp"
#include "afr_sst_point.hpp"
#include "afr_step.hpp"

using namespace afr;

void afr_set_error(const std::string& msg);

namespace {

constexpr int kThreads = 256;
constexpr double kCip = 4.0;   // interior-penalty coefficient, viscous_flux.py:183-192

// primitive state of cell c; off: the conserved offset of the stored state (null: absolute)
template <typename real>
__device__ __forceinline__ void prim_at(const real* U, int64_t c, const double* off, double q[5]) {
    double u[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) u[v] = off ? (double)U[c * 5 + v] + off[v] : (double)U[c * 5 + v];
    cons2prim(u, q);
}

template <typename real>
__global__ void __launch_bounds__(kThreads) k_p0_residual(AfrP0Desc d, const real* __restrict__ U,
                                                          const real* __restrict__ mu_t, real* __restrict__ R,
                                                          double mu) {
    const real* det = (const real*)d.det;
    const real* adj = (const real*)d.adj;
    const double* vol = d.cell_volume;
    for (int64_t c = (int64_t)blockIdx.x * kThreads + threadIdx.x; c < d.n_cells;
         c += (int64_t)gridDim.x * kThreads) {
        double qc[5];
        prim_at(U, c, d.state_ref, qc);
        double res[5] = {0, 0, 0, 0, 0};
        // the ghost state the cell keeps: that of its last boundary record
        double qg[5];
        const int gr = d.cell_ghost_rec[c];
        if (gr >= 0) {
            const double* gg = d.rec_geom + (int64_t)gr * 4;
            const double n[3] = {gg[0], gg[1], gg[2]};
            ghost_state(d.rec_bc[gr], qc, n, d.bc_params, qg);
        }
        const double iv = 1.0 / vol[c];
        for (int k = 0; k < 8; ++k) {
            const int e = d.cell_inc[c * 8 + k];
            if (e < 0) break;
            const int r = e >> 4;
            const int role = (e >> 3) & 1;
            const double* gg = d.rec_geom + (int64_t)r * 4;
            const double n[3] = {gg[0], gg[1], gg[2]};
            const int nb = d.rec_nb[r];
            double qo[5], qn[5], f[5];
            if (role == 0) {
#pragma unroll
                for (int v = 0; v < 5; ++v) qo[v] = qc[v];
                if (nb >= 0) prim_at(U, nb, d.state_ref, qn);
                else {
#pragma unroll
                    for (int v = 0; v < 5; ++v) qn[v] = qg[v];
                }
            } else {
                prim_at(U, d.rec_owner[r], d.state_ref, qo);
#pragma unroll
                for (int v = 0; v < 5; ++v) qn[v] = qc[v];
            }
            ausm_up(qo, qn, n, f);
            const double s = (role == 0 ? -gg[3] : gg[3]) * iv;
#pragma unroll
            for (int v = 0; v < 5; ++v) res[v] += f[v] * s;
            // the boundary interior penalty of the owner-primary boundary records
            const int fl = d.rec_flags[r];
            if (role == 0 && nb < 0 && (fl & 1)) {
                double qb[5];
                ghost_state(d.rec_bc[r], qc, n, d.bc_params, qb);
                const int axis = (fl >> 1) & 3;
                const double side = (fl & 8) ? 1.0 : -1.0;
                const real* ar = adj + c * 9 + axis * 3;
                const double a0 = (double)ar[0], a1 = (double)ar[1], a2 = (double)ar[2];
                const double mag = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
                const double dj = (double)det[c];
                double h = cbrt(fabs(dj));
                h = h > 1e-300 ? h : 1e-300;
                const double eta = kCip * (mu + (mu_t ? (double)mu_t[c] : 0.0)) / h;
                const double gp = side < 0 ? -0.5 : 0.5;
#pragma unroll
                for (int v = 1; v < 4; ++v) res[v] += gp * (-eta * mag * side * (qc[v] - qb[v])) / dj;
            }
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) R[c * 5 + v] = (real)res[v];
    }
}

template <typename real>
__global__ void __launch_bounds__(kThreads) k_p0_stage(const real* U0, const real* Ua, const real* __restrict__ R,
                                                       const real* __restrict__ dt, real* U_out, int64_t n,
                                                       const double* off, double c0, double c1, double c2,
                                                       double p_floor) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const double d = (double)dt[i];
        double q[5], qa[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            double val = c0 * (double)U0[i * 5 + v] + c2 * d * (double)R[i * 5 + v];
            if (c1 != 0.0) val += c1 * (double)Ua[i * 5 + v];
            q[v] = (double)(real)val;
            qa[v] = off ? q[v] + off[v] : q[v];
        }
        // positivity (src/autoflowcfd/core/time_integration/positivity.py:13-45) on the absolute
        // state; a projected component is stored as its value less the offset, the others as they were
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
        const double o0 = off ? off[0] : 0.0, o1 = off ? off[1] : 0.0, o2 = off ? off[2] : 0.0;
        const double o3 = off ? off[3] : 0.0, o4 = off ? off[4] : 0.0;
        U_out[i * 5 + 0] = low_rho ? (real)(rho - o0) : (real)q[0];
        U_out[i * 5 + 1] = fast ? (real)(m1 - o1) : (real)q[1];
        U_out[i * 5 + 2] = fast ? (real)(m2 - o2) : (real)q[2];
        U_out[i * 5 + 3] = fast ? (real)(m3 - o3) : (real)q[3];
        U_out[i * 5 + 4] = low_p ? (real)(E - o4) : (real)q[4];
    }
}

template <typename real>
__global__ void __launch_bounds__(kThreads) k_interp_order(const real* __restrict__ X, real* __restrict__ Y,
                                                           const double* __restrict__ W, int64_t n_cells, int ns_old,
                                                           int ns_new, int nv) {
    const int64_t n = n_cells * ns_new * nv;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const int64_t c = i / ((int64_t)ns_new * nv);
        const int a = (int)((i / nv) % ns_new), v = (int)(i % nv);
        double acc = 0.0;
        for (int b = 0; b < ns_old; ++b) acc += W[a * ns_old + b] * (double)X[(c * ns_old + b) * nv + v];
        Y[i] = (real)acc;
    }
}


using afr_sst::A1;
using afr_sst::B1;
using afr_sst::B2;
using afr_sst::BSTAR;
using afr_sst::KAP;
using afr_sst::SK1;
using afr_sst::SK2;
using afr_sst::SW1;
using afr_sst::SW2;
using afr_sst::dmax;
using afr_sst::dmin;
using afr_sst::finite_or;

// work_sst per cell: 0 Gamma_k, 1 Gamma_omega, 2 S_k / rho, 3 S_omega / rho, 4 beta, 5 nu_t, 6 T_k, 7 T_omega
constexpr int kSstW = 8;

template <typename real>
__global__ void __launch_bounds__(kThreads) k_p0_sst_point(AfrP0Desc d, const real* __restrict__ U,
                                                           const real* __restrict__ k, const real* __restrict__ w,
                                                           double mu) {
    const real* dw = (const real*)d.d_wall;
    for (int64_t c = (int64_t)blockIdx.x * kThreads + threadIdx.x; c < d.n_cells;
         c += (int64_t)gridDim.x * kThreads) {
        const double rho = dmax((double)U[c * 5] + (d.state_ref ? d.state_ref[0] : 0.0), 1e-10);
        const double kc = (double)k[c], wc = (double)w[c], dc = (double)dw[c];
        // sources with zero gradients (S = 0, grad k . grad omega = 0), mu = 1.8e-5 as the reference's
        const double musrc = 1.8e-5;
        const double nu = musrc / rho;
        const double ks = dmin(kc, 1e40), wsf = dmax(dmin(wc, 1e40), 1e-10);
        const double CD = dmax(2.0 * rho * SW2 / wsf * 0.0, 1e-10);
        const double F1 = afr_sst::blend_F1(ks, wsf, dc, nu, rho, CD);
        const double F2 = afr_sst::blend_F2(ks, wsf, dc, nu);
        const double beta = F1 * B1 + (1.0 - F1) * B2;
        double nut = A1 * dmax(ks, 1e-10) / dmax(A1 * dmax(wsf, 1e-10), F2 * 1e-10);
        nut = dmin(nut, 1.0e5 * musrc / rho);
        nut = isfinite(nut) ? nut : 0.0;
        const double Pk = dmin(nut * rho * 0.0, 10.0 * BSTAR * rho * ks * wsf);
        const double Sk = finite_or(Pk - rho * BSTAR * ks * wsf, 0.0);
        const double g1 = B1 / BSTAR - SW1 * KAP * KAP / sqrt(BSTAR);
        const double g2 = B2 / BSTAR - SW2 * KAP * KAP / sqrt(BSTAR);
        const double gam = F1 * g1 + (1.0 - F1) * g2;
        const double Sw = finite_or(rho * gam * 0.0 - rho * beta * wsf * wsf + 2.0 * rho * (1.0 - F1) * SW2 / wsf * 0.0,
                                    0.0);
        // diffusivities with the transport's own F1 (molecular mu, raw k / omega)
        const double CDt = dmax(2.0 * rho * SW2 / dmax(wc, 1e-10) * 0.0, 1e-10);
        const double F1t = afr_sst::blend_F1(kc, wc, dc, mu / rho, rho, CDt);
        const double rnut = rho * nut;
        double* o = d.work_sst + c * kSstW;
        o[0] = mu + (F1t * SK1 + (1.0 - F1t) * SK2) * rnut;
        o[1] = mu + (F1t * SW1 + (1.0 - F1t) * SW2) * rnut;
        o[2] = Sk / rho;
        o[3] = Sw / rho;
        o[4] = beta;
        o[5] = nut;
    }
}

template <typename real>
__global__ void __launch_bounds__(kThreads) k_p0_sst_transport(AfrP0Desc d, const real* __restrict__ U,
                                                               const real* __restrict__ k,
                                                               const real* __restrict__ w) {
    const real* det = (const real*)d.det;
    for (int64_t c = (int64_t)blockIdx.x * kThreads + threadIdx.x; c < d.n_cells;
         c += (int64_t)gridDim.x * kThreads) {
        double conv[2] = {0.0, 0.0}, diff[2] = {0.0, 0.0};
        const double dj = (double)det[c];
        for (int q = 0; q < 8; ++q) {
            const int e = d.cell_inc[c * 8 + q];
            if (e < 0) break;
            const int r = e >> 4, role = (e >> 3) & 1, code = e & 7;
            const double g = (code & 1) ? 0.5 : -0.5;   // g'(0) of the side's Radau correction
            const int64_t o = d.rec_owner[r];
            double uo[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) uo[v] = d.state_ref ? (double)U[o * 5 + v] + d.state_ref[v] : (double)U[o * 5 + v];
            const double ro = dmax(uo[0], 1e-10);
            const double* n = d.rec_geom + (int64_t)r * 4;
            const double vo[3] = {uo[1] / ro, uo[2] / ro, uo[3] / ro};
            const double mass = ro * vo[0] * n[0] + ro * vo[1] * n[1] + ro * vo[2] * n[2];
            const int s0 = d.rec_src[r * 2], s1 = d.rec_src[r * 2 + 1];
            const double w0 = d.rec_w[r * 2], w1 = d.rec_w[r * 2 + 1];
            const bool none = s0 < 0 && s1 < 0;
            const bool wall = (d.rec_flags[r] >> 4) & 1;
            for (int f = 0; f < 2; ++f) {
                const real* phi = f == 0 ? k : w;
                const double po = (double)phi[o];
                const double go = d.work_sst[o * kSstW + f];
                double pn, gn;
                if (none) {
                    pn = (f == 0 && wall) ? -po : po;
                    gn = go;
                } else {
                    pn = s0 >= 0 ? w0 * (double)phi[s0] : 0.0;
                    gn = s0 >= 0 ? w0 * d.work_sst[(int64_t)s0 * kSstW + f] : 0.0;
                    if (s1 >= 0) {
                        pn += w1 * (double)phi[s1];
                        gn += w1 * d.work_sst[(int64_t)s1 * kSstW + f];
                    }
                }
                const double up = mass >= 0.0 ? po : pn;
                const double cc = mass * (up - po);
                const double cd = 0.5 * (go + gn) * (0.5 * (po + pn) - po);
                if (role == 0) {
                    conv[f] += -(g * cc) / dj;
                    diff[f] += -(g * cd) / dj;
                } else {
                    conv[f] += (g * cc) / dj;
                    diff[f] += (g * cd) / dj;
                }
            }
        }
        const double rho = dmax((double)U[c * 5] + (d.state_ref ? d.state_ref[0] : 0.0), 1e-10);
        for (int f = 0; f < 2; ++f) d.work_sst[c * kSstW + 6 + f] = finite_or((conv[f] - diff[f]) / rho, 0.0);
    }
}

template <typename real>
__global__ void __launch_bounds__(kThreads) k_p0_sst_update(AfrP0Desc d, const real* __restrict__ U, real* k, real* w,
                                                            real* nu_t, real* mu_t, const real* __restrict__ dt) {
    for (int64_t c = (int64_t)blockIdx.x * kThreads + threadIdx.x; c < d.n_cells;
         c += (int64_t)gridDim.x * kThreads) {
        const double* o = d.work_sst + c * kSstW;
        const double h = (double)dt[c];
        const double kc = (double)k[c], wc = (double)w[c];
        const double ws = dmax(wc, 1e-10);
        double dk = finite_or(o[2] / (1.0 + h * BSTAR * ws), 0.0) + o[6];
        double dw = finite_or(o[3] / (1.0 + h * o[4] * ws), 0.0) + o[7];
        dk = finite_or(dk, 0.0);
        dw = finite_or(dw, 0.0);
        double kn = kc + h * dk, wn = wc + h * dw;
        kn = isfinite(kn) ? kn : 1e-12;
        wn = isfinite(wn) ? wn : 1e-12;
        k[c] = (real)dmax(kn, 1e-12);
        w[c] = (real)dmax(wn, 1e-12);
        const double rho = dmax((double)U[c * 5] + (d.state_ref ? d.state_ref[0] : 0.0), 1e-10);
        nu_t[c] = (real)o[5];
        mu_t[c] = (real)(rho * o[5]);
    }
}

// the P0 state's storage for the shared local-time-step kernel: the offset row on the device, or none
StateOff p0_state_off(const AfrP0Desc& d) {
    StateOff so{};
    so.table = d.state_ref;
    so.mode = d.state_ref ? 2 : 0;
    return so;
}

int grid_of(int64_t n) {
    int64_t b = (n + kThreads - 1) / kThreads;
    return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

bool desc_ok(const AfrP0Desc* d) {
    return d && d->abi_version == AFR_ABI_VERSION && (d->dtype == AFR_F32 || d->dtype == AFR_F64) && d->n_cells >= 0 &&
           d->det && d->adj && d->mfs && d->vol1 && d->cell_volume && d->rec_owner && d->rec_nb && d->rec_bc &&
           d->rec_flags && d->rec_geom && d->cell_inc && d->cell_ghost_rec && d->bc_params;
}

template <typename real>
int residual_t(const AfrP0Desc& d, const void* U, const void* mu_t, void* R, double mu, hipStream_t st) {
    hipLaunchKernelGGL((k_p0_residual<real>), dim3(grid_of(d.n_cells)), dim3(kThreads), 0, st, d, (const real*)U,
                       (const real*)mu_t, (real*)R, mu);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename real>
int stage_t(const void* U0, const void* Ua, const void* R, const void* dt, void* out, int64_t n, const double* off,
            double c0, double c1, double c2, hipStream_t st) {
    hipLaunchKernelGGL((k_p0_stage<real>), dim3(grid_of(n)), dim3(kThreads), 0, st, (const real*)U0,
                       (const real*)(Ua ? Ua : U0), (const real*)R, (const real*)dt, (real*)out, n, off, c0, c1, c2,
                       1.0);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// the residual norm and the three Shu-Osher stages of U with the local time step already in work_dt
template <typename real>
int stages_t(const AfrP0Desc& d, void* U, const void* mu_t, double mu, int n_vars_total, double* norm,
             hipStream_t st) {
    const int64_t C = d.n_cells;
    real* u = (real*)U;
    real* r = (real*)d.work_r;
    real* s1 = (real*)d.work_s1;
    real* s2 = (real*)d.work_s2;
    real* dt = (real*)d.work_dt;
    if (residual_t<real>(d, u, mu_t, r, mu, st)) return -1;
    if (afr_step::launch_norm<real>(r, C * 5, (double)C * n_vars_total, d.work_part, afr_step::kNormParts, norm, st))
        return -1;
    if (stage_t<real>(u, nullptr, r, dt, s1, C, d.state_ref, 1.0, 0.0, 1.0, st)) return -1;
    if (residual_t<real>(d, s1, mu_t, r, mu, st)) return -1;
    if (stage_t<real>(u, s1, r, dt, s2, C, d.state_ref, 0.75, 0.25, 0.25, st)) return -1;
    if (residual_t<real>(d, s2, mu_t, r, mu, st)) return -1;
    return stage_t<real>(u, s2, r, dt, u, C, d.state_ref, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, st);
}

template <typename real>
int step_t(const AfrP0Desc& d, void* U, const void* mu_t, double mu, int n_vars_total, double* norm, hipStream_t st) {
    if (afr_step::launch_dt_local<real>((const real*)U, p0_state_off(d), (const real*)mu_t,
                                        (const real*)d.det, (const real*)d.mfs,
                                        (const real*)d.vol1, (real*)d.work_dt, d.n_cells, 1, mu, 0, 0, st))
        return -1;
    return stages_t<real>(d, U, mu_t, mu, n_vars_total, norm, st);
}

template <typename real>
int sst_step_t(const AfrP0Desc& d, void* U, void* k, void* w, void* nu_t, void* mu_t, double mu, int n_vars_total,
               double* norm, hipStream_t st) {
    const int g = grid_of(d.n_cells);
    real* dt = (real*)d.work_dt;
    // local time step with the previous step's eddy viscosity (nu_t held, mu_eff = mu + rho nu_t)
    if (afr_step::launch_dt_local<real>((const real*)U, p0_state_off(d), (const real*)nu_t,
                                        (const real*)d.det, (const real*)d.mfs,
                                        (const real*)d.vol1, dt, d.n_cells, 1, mu, 0, 1, st))
        return -1;
    hipLaunchKernelGGL((k_p0_sst_point<real>), dim3(g), dim3(kThreads), 0, st, d, (const real*)U, (const real*)k,
                       (const real*)w, mu);
    hipLaunchKernelGGL((k_p0_sst_transport<real>), dim3(g), dim3(kThreads), 0, st, d, (const real*)U,
                       (const real*)k, (const real*)w);
    hipLaunchKernelGGL((k_p0_sst_update<real>), dim3(g), dim3(kThreads), 0, st, d, (const real*)U, (real*)k,
                       (real*)w, (real*)nu_t, (real*)mu_t, (const real*)dt);
    if (hipGetLastError() != hipSuccess) return -1;
    return stages_t<real>(d, U, mu_t, mu, n_vars_total, norm, st);
}

}  // namespace

extern "C" {

int afr_p0_residual(const AfrP0Desc* d, const void* U, const void* mu_t, double mu, void* R, void* stream) {
    if (!desc_ok(d) || !U || !R) {
        afr_set_error("afr_p0_residual: bad descriptor or null argument");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    const int rc = d->dtype == AFR_F64 ? residual_t<double>(*d, U, mu_t, R, mu, st)
                                       : residual_t<float>(*d, U, mu_t, R, mu, st);
    if (rc) afr_set_error("afr_p0_residual: launch failed");
    return rc;
}

int afr_p0_ssp_rk3_step(const AfrP0Desc* d, void* U, const void* mu_t, double mu, int32_t n_vars_total, double* norm,
                        void* stream) {
    if (!desc_ok(d) || !U || !norm || !d->work_r || !d->work_s1 || !d->work_s2 || !d->work_dt || !d->work_part ||
        n_vars_total < 5) {
        afr_set_error("afr_p0_ssp_rk3_step: bad descriptor, work buffer or argument");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    const int rc = d->dtype == AFR_F64 ? step_t<double>(*d, U, mu_t, mu, n_vars_total, norm, st)
                                       : step_t<float>(*d, U, mu_t, mu, n_vars_total, norm, st);
    if (rc) afr_set_error("afr_p0_ssp_rk3_step: launch failed");
    return rc;
}

int afr_p0_sst_step(const AfrP0Desc* d, void* U, void* k, void* omega, void* nu_t, void* mu_t, double mu,
                    int32_t n_vars_total, double* norm, void* stream) {
    if (!desc_ok(d) || !U || !k || !omega || !nu_t || !mu_t || !norm || !d->work_r || !d->work_s1 || !d->work_s2 ||
        !d->work_dt || !d->work_part || !d->rec_src || !d->rec_w || !d->d_wall || !d->work_sst || n_vars_total < 5) {
        afr_set_error("afr_p0_sst_step: bad descriptor (SST arrays), work buffer or argument");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    const int rc = d->dtype == AFR_F64 ? sst_step_t<double>(*d, U, k, omega, nu_t, mu_t, mu, n_vars_total, norm, st)
                                       : sst_step_t<float>(*d, U, k, omega, nu_t, mu_t, mu, n_vars_total, norm, st);
    if (rc) afr_set_error("afr_p0_sst_step: launch failed");
    return rc;
}

int afr_interpolate_order(const void* X_old, void* X_new, const double* W, int64_t n_cells, int32_t ns_old,
                          int32_t ns_new, int32_t n_vars, int32_t dtype, void* stream) {
    if (!X_old || !X_new || !W || n_cells < 0 || ns_old < 1 || ns_new < 1 || ns_old > 64 || ns_new > 64 ||
        n_vars < 1 || (dtype != AFR_F32 && dtype != AFR_F64)) {
        afr_set_error("afr_interpolate_order: bad argument");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = n_cells * ns_new * n_vars;
    if (dtype == AFR_F64)
        hipLaunchKernelGGL((k_interp_order<double>), dim3(grid_of(n)), dim3(kThreads), 0, st, (const double*)X_old,
                           (double*)X_new, W, n_cells, ns_old, ns_new, n_vars);
    else
        hipLaunchKernelGGL((k_interp_order<float>), dim3(grid_of(n)), dim3(kThreads), 0, st, (const float*)X_old,
                           (float*)X_new, W, n_cells, ns_old, ns_new, n_vars);
    if (hipGetLastError() != hipSuccess) {
        afr_set_error("afr_interpolate_order: launch failed");
        return -1;
    }
    return 0;
}

}  // extern "C"
