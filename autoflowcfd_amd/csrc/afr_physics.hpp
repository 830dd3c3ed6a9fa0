// Point physics of the FR residual path (device side, templated on precision).
//
// Each function restates one reference routine; the branch structure and the
// floors are the reference's own:
//   euler_flux      src/autoflowcfd/core/fr_operators/flux_kernels.py:29-68
//   ausm_up         src/autoflowcfd/core/fr_operators/kernels.py:20-155
//   ghost_state     src/autoflowcfd/boundary/fr_ghost_state.py:26-189
//   viscous_flux    src/autoflowcfd/core/fr_operators/flux_kernels.py:72-137
#pragma once
#include <hip/hip_runtime.h>

// Occupancy targets (waves per SIMD) of the heavy kernels, as build parameters: 0 leaves the
// register allocation to the compiler; n > 0 asks for at least n waves per SIMD (the compiler
// then budgets registers for it, spilling if it must).  Measured on the 82^3 P2 SST residual:
// the interface kernel 3.52 (7 waves, 28 B/lane spilled), 3.22 (6 waves, no spill), 3.55 ms
// (compiler's choice, 5 waves); the SST update 3.90 (8 waves, 56 B/lane spilled), 3.83 (7 waves,
// 20 B), 3.88 ms (6 waves, no spill) for the whole SST phase; the SST cell, inviscid and viscous
// volume kernels are bounded by LDS or already at their register limit.
#ifndef AFR_WPE_SST_CELL
#define AFR_WPE_SST_CELL 0
#endif
#ifndef AFR_WPE_SST_UPDATE
#define AFR_WPE_SST_UPDATE 7
#endif
#ifndef AFR_WPE_VISC_IFACE
#define AFR_WPE_VISC_IFACE 6
#endif
#ifndef AFR_WPE_VISC_VOLUME
#define AFR_WPE_VISC_VOLUME 0
#endif
#ifndef AFR_WPE_PRISM
#define AFR_WPE_PRISM 0
#endif
#define AFR_WPE_ATTR_0
#define AFR_WPE_ATTR_N(n) __attribute__((amdgpu_waves_per_eu(n, 8)))
#define AFR_WPE_SEL(n) AFR_WPE_SEL2(n)
#define AFR_WPE_SEL2(n) AFR_WPE_ATTR_##n
#define AFR_WPE_ATTR_4 AFR_WPE_ATTR_N(4)
#define AFR_WPE_ATTR_5 AFR_WPE_ATTR_N(5)
#define AFR_WPE_ATTR_6 AFR_WPE_ATTR_N(6)
#define AFR_WPE_ATTR_7 AFR_WPE_ATTR_N(7)
#define AFR_WPE_ATTR_8 AFR_WPE_ATTR_N(8)

namespace afr {

template <typename real> struct Tiny;
template <> struct Tiny<float> { static constexpr float v = 1.17549435e-38f; };
template <> struct Tiny<double> { static constexpr double v = 1e-300; };

constexpr double kGamma = 1.4;
constexpr double kRair = 287.0;

enum BcCode { BC_INTERIOR = -1, BC_WALL = 0, BC_FARFIELD = 1, BC_INLET = 2, BC_OUTLET = 3,
              BC_SYMMETRY = 4, BC_SLIP_WALL = 5, BC_EXTRAP = 6, BC_MOVING_WALL = 7 };

template <typename real>
__host__ __device__ __forceinline__ real rmax(real a, real b) { return a > b ? a : b; }

// 1/x for positive normal x.  Device: hardware reciprocal estimate refined by
// Newton steps (two for FP64, one for FP32), within a few ulp of the correctly
// rounded quotient; host: IEEE division.
__host__ __device__ __forceinline__ double recip(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
#else
    return 1.0 / x;
#endif
}
__host__ __device__ __forceinline__ float recip(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    float r = __builtin_amdgcn_rcpf(x);
    return fmaf(r, fmaf(-x, r, 1.0f), r);
#else
    return 1.0f / x;
#endif
}
template <typename real>
__host__ __device__ __forceinline__ real rmin(real a, real b) { return a < b ? a : b; }

// Stored state -> conserved state.  A mesh may store every state as the offset from a per-case
// conserved reference (the case's free stream): the kernels add the offset back in FP64 when they
// load a point.  A single-case mesh carries its offset by value (a kernel argument, held in scalar
// registers); a batched mesh reads the row of the cell's case from a device table.  Without an
// offset the storage holds the state itself and the load is a plain widening (bit for bit the same
// values as before the offset existed).
typedef const __attribute__((address_space(4))) int* ConstIntPtr;       // read-only for the kernel:
typedef const __attribute__((address_space(4))) double* ConstDblPtr;    // uniform loads allowed

__device__ __forceinline__ int64_t lane_value(int64_t x, int lane) {
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)x, lane);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), lane);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

struct StateOff {
    const double* table;   // mode 2: (n_cases, 5) FP64 device table
    const int* cell_case;  // mode 2: the case of every cell (null: row 0)
    double one[5];         // mode 1: the offset
    int mode;              // 0 absolute storage, 1 one offset by value, 2 per-case table
    int sorted;            // mode 2: cell_case is non-decreasing on [0, split) and on [split, n_cells)
    int64_t split;         // first cell of the second ascending range (the first tet of a batch)

    // The offset of the case of `cell` (lanes may hold different cells; call it where every lane
    // of the wave that needs a row is active).  With sorted cell_case, a wave whose active lanes
    // all hold cells between those of its first and last active lanes, both on one side of
    // `split` and of one case, takes that case's row with wave-uniform loads; otherwise every
    // lane looks up its own case and row.  Both give the same values.
    __device__ __forceinline__ void row(int64_t cell, double off[5]) const {
        if (mode == 1) {
#pragma unroll
            for (int v = 0; v < 5; ++v) off[v] = one[v];
            return;
        }
        if (mode != 2) {
#pragma unroll
            for (int v = 0; v < 5; ++v) off[v] = 0.0;
            return;
        }
        if (!cell_case) {
#pragma unroll
            for (int v = 0; v < 5; ++v) off[v] = ((ConstDblPtr)table)[v];
            return;
        }
        bool wave_row = false;
        if (sorted) {
            const uint64_t act = __builtin_amdgcn_read_exec();
            const int64_t a = lane_value(cell, __builtin_ctzll(act));
            const int64_t b = lane_value(cell, 63 - __builtin_clzll(act));
            const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
            if ((lo < split) == (hi < split) && __all(lo <= cell && cell <= hi)) {
                const int ka = ((ConstIntPtr)cell_case)[lo], kb = ((ConstIntPtr)cell_case)[hi];
                if (ka == kb) {
                    const ConstDblPtr r = (ConstDblPtr)table + 5 * (int64_t)ka;
#pragma unroll
                    for (int v = 0; v < 5; ++v) off[v] = r[v];
                    wave_row = true;
                }
            }
        }
        if (!wave_row) {
            const double* r = table + 5 * (int64_t)cell_case[cell];
#pragma unroll
            for (int v = 0; v < 5; ++v) off[v] = r[v];
        }
    }
};

template <typename real>
__device__ __forceinline__ void load_state(const real* U, int64_t i, const StateOff& so, int64_t cell, double u[5]) {
#pragma unroll
    for (int v = 0; v < 5; ++v) u[v] = (double)U[i * 5 + v];
    if (so.mode) {
        double off[5];
        so.row(cell, off);
#pragma unroll
        for (int v = 0; v < 5; ++v) u[v] += off[v];
    }
}

// conserved (rho, rho u, rho v, rho w, rho E) -> primitive (rho, u, v, w, p); rho floor 1e-10, no p floor
template <typename real>
__host__ __device__ __forceinline__ void cons2prim(const real U[5], real Q[5]) {
    const real g1 = real(kGamma - 1.0);
    real rho = rmax(U[0], real(1e-10));
    const real ir = recip(rho);   // one reciprocal instead of four divisions
    real u = U[1] * ir, v = U[2] * ir, w = U[3] * ir, E = U[4] * ir;
    Q[0] = rho; Q[1] = u; Q[2] = v; Q[3] = w;
    Q[4] = g1 * rho * (E - real(0.5) * (u * u + v * v + w * w));
}

// physical flux F[i][v] of primitive state q
template <typename real>
__host__ __device__ __forceinline__ void euler_flux(const real q[5], real F[3][5]) {
    const real g1 = real(kGamma - 1.0);
    real rho = q[0], u = q[1], v = q[2], w = q[3], p = q[4];
    real irs = recip(rmax(rho, real(1e-10)));
    real ke = real(0.5) * (u * u + v * v + w * w);
    real rhoE = rho * (p * real(1.0 / (kGamma - 1.0)) * irs + ke);
    real H = (rhoE + p) * irs;
    real rH = rho * H;
    real vel[3] = {u, v, w};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        real mf = rho * vel[i];
        F[i][0] = mf;
        F[i][1] = mf * u + (i == 0 ? p : real(0));
        F[i][2] = mf * v + (i == 1 ? p : real(0));
        F[i][3] = mf * w + (i == 2 ? p : real(0));
        F[i][4] = rH * vel[i];
    }
}

// sqrt(x) for x >= 0.  Device: hardware reciprocal square root estimate refined by
// one Newton step and a final residual correction (within about 1 ulp); host: sqrt.
__host__ __device__ __forceinline__ double sqrt_r(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rsq(x);
    r = r * fma(-0.5 * x * r, r, 1.5);
    double s = x * r;
    s = fma(fma(-s, s, x), 0.5 * r, s);
    return x > 0.0 ? s : 0.0;
#else
    return sqrt(x);
#endif
}
__host__ __device__ __forceinline__ float sqrt_r(float x) { return sqrtf(x); }

// directional flux sum_i n_i F[i][v] of primitive state q (n need not be unit)
template <typename real>
__host__ __device__ __forceinline__ void euler_flux_dir(const real q[5], const real n[3], real F[5]) {
    real rho = q[0], u = q[1], v = q[2], w = q[3], p = q[4];
    real irs = recip(rmax(rho, real(1e-10)));
    real ke = real(0.5) * (u * u + v * v + w * w);
    real rhoE = rho * (p * real(1.0 / (kGamma - 1.0)) * irs + ke);
    real rH = rho * ((rhoE + p) * irs);
    real un = u * n[0] + v * n[1] + w * n[2];
    real mf = rho * un;
    F[0] = mf;
    F[1] = mf * u + p * n[0];
    F[2] = mf * v + p * n[1];
    F[3] = mf * w + p * n[2];
    F[4] = rH * un;
}

// AUSM+up normal flux density for unit direction n
template <typename real>
__host__ __device__ __forceinline__ void ausm_up(const real qL[5], const real qR[5], const real n[3], real out[5]) {
    const real g = real(1.4), alpha = real(0.1875), beta = real(0.5);
    real rhoL = rmax(qL[0], real(1e-6)), rhoR = rmax(qR[0], real(1e-6));
    real pL = rmax(qL[4], real(10.0)), pR = rmax(qR[4], real(10.0));
    real uL = qL[1], vL = qL[2], wL = qL[3];
    real uR = qR[1], vR = qR[2], wR = qR[3];
    real unL = uL * n[0] + vL * n[1] + wL * n[2];
    real unR = uR * n[0] + vR * n[1] + wR * n[2];
    real irhoL = recip(rhoL), irhoR = recip(rhoR);
    real aL = sqrt_r(rmax(g * pL * irhoL, real(1e-10)));
    real aR = sqrt_r(rmax(g * pR * irhoR, real(1e-10)));
    real ML = unL * recip(rmax(aL, real(1e-10)));
    real MR = unR * recip(rmax(aR, real(1e-10)));
    real ah = real(0.5) * (aL + aR);
    real rh = real(0.5) * (rhoL + rhoR);
    real iah2 = recip(ah * ah);
    real Mbar2 = (unL * unL + unR * unR) * real(0.5) * iah2;
    real M0sq = rmin(real(1.0), rmax(Mbar2, real(0.01)));
    real M0 = sqrt_r(M0sq);
    real fa = rmax(M0 * (real(2.0) - M0), real(1e-6));
    real Mp, Mm, Pp, Pm;
    {
        real M = ML;
        real aM = M < 0 ? -M : M;
        if (aM >= real(1)) {
            Mp = real(0.5) * (M + aM);
            Pp = real(0.5) * (real(1) + (M > 0 ? real(1) : (M < 0 ? real(-1) : real(0))));
        } else {
            real t = M * M - real(1);
            Mp = real(0.25) * (M + real(1)) * (M + real(1)) + alpha * t * t;
            Pp = real(0.25) * ((M + real(1)) * (M + real(1)) * (real(2) - M) + beta * M * t * t);
        }
    }
    {
        real M = MR;
        real aM = M < 0 ? -M : M;
        if (aM >= real(1)) {
            Mm = real(0.5) * (M - aM);
            Pm = real(0.5) * (real(1) - (M > 0 ? real(1) : (M < 0 ? real(-1) : real(0))));
        } else {
            real t = M * M - real(1);
            Mm = real(-0.25) * (M - real(1)) * (M - real(1)) - alpha * t * t;
            Pm = real(0.25) * ((M - real(1)) * (M - real(1)) * (real(2) + M) - beta * M * t * t);
        }
    }
    const real Kp = real(0.25), Ku = real(0.75);
    real Mh = Mp + Mm;
    real Mpress = -Kp * rmax(real(1.0) - Mbar2, real(0.0)) * (pR - pL) * recip(fa * rh) * iah2;
    real mass = real(0.5) * (rhoL * aL + rhoR * aR) * (Mh + Mpress);
    real ph = Pp * pL + Pm * pR - Ku * Pp * Pm * (rhoL + rhoR) * fa * ah * (unR - unL);
    bool up = mass >= real(0);
    out[0] = mass;
    out[1] = mass * (up ? uL : uR) + ph * n[0];
    out[2] = mass * (up ? vL : vR) + ph * n[1];
    out[3] = mass * (up ? wL : wR) + ph * n[2];
    const real gg = g / (g - real(1));
    real hL = gg * pL * irhoL + real(0.5) * (uL * uL + vL * vL + wL * wL);
    real hR = gg * pR * irhoR + real(0.5) * (uR * uR + vR * vR + wR * wR);
    out[4] = mass * (up ? hL : hR);
}

// primitive ghost state of a boundary slot (src/autoflowcfd/boundary/fr_ghost_state.py:26-189).
// bc = type | (row << 4): the boundary type (BcCode) and the row of the (8, 5) table bcp that holds
// its reference state, so that boundary groups of one type may carry states of their own
// (FARFIELD / INLET: the state; OUTLET: the pressure in column 4; MOVING_WALL: the wall velocity in
// columns 1-3, ghost velocity 2 v_wall - v as fr_ghost_state.py:26-53 forms it).
template <typename real>
__host__ __device__ __forceinline__ void ghost_state(int bc, const real qi[5], const real n[3], const double* bcp,
                                                     real qg[5]) {
#pragma unroll
    for (int v = 0; v < 5; ++v) qg[v] = qi[v];
    const double* p = bcp + 5 * (bc >> 4);
    real vn = qi[1] * n[0] + qi[2] * n[1] + qi[3] * n[2];
    switch (bc & 15) {
        case BC_WALL:
            qg[1] = -qi[1]; qg[2] = -qi[2]; qg[3] = -qi[3];
            break;
        case BC_MOVING_WALL:
#pragma unroll
            for (int v = 1; v < 4; ++v) qg[v] = real(2) * (real)p[v] - qi[v];
            break;
        case BC_SLIP_WALL:
        case BC_SYMMETRY:
            qg[1] = qi[1] - real(2) * vn * n[0];
            qg[2] = qi[2] - real(2) * vn * n[1];
            qg[3] = qi[3] - real(2) * vn * n[2];
            break;
        case BC_FARFIELD:
#pragma unroll
            for (int v = 0; v < 5; ++v) qg[v] = (real)p[v];
            break;
        case BC_INLET:
            if (vn < real(0)) {
#pragma unroll
                for (int v = 0; v < 5; ++v) qg[v] = (real)p[v];
            }
            break;
        case BC_OUTLET:
            if (vn > real(0)) qg[4] = (real)p[4];
            break;
        default:  // BC_EXTRAP and anything else: zero-gradient copy
            break;
    }
}

// viscous flux G[i][v]; gv[i][j] = du_i/dx_j
template <typename real>
__host__ __device__ __forceinline__ void viscous_flux(const real q[5], const real gv[3][3], const real gT[3], real mu, real Pr,
                                             real mut, real Prt, real G[3][5]) {
    real mt = mu + mut;
    real div = gv[0][0] + gv[1][1] + gv[2][2];
    real lam = real(-2.0 / 3.0) * mt;
    real tau[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            tau[i][j] = real(2) * mt * (real(0.5) * (gv[i][j] + gv[j][i])) + (i == j ? lam * div : real(0));
    const real cp = real(kGamma * kRair / (kGamma - 1.0));
    real k = mu * cp / Pr + mut * cp / Prt;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        G[i][0] = real(0);
        G[i][1] = tau[i][0];
        G[i][2] = tau[i][1];
        G[i][3] = tau[i][2];
        real work = q[1] * tau[0][i] + q[2] * tau[1][i] + q[3] * tau[2][i];
        G[i][4] = work - k * gT[i];
    }
}

// A . G for the viscous flux G of viscous_flux, contracted with a direction A (3): the five
// components A_i G[i][v], computed from tau . A (tau is symmetric) without forming G.
template <typename real>
__host__ __device__ __forceinline__ void viscous_flux_dot(const real q[5], const real gv[3][3], const real gT[3],
                                                         real mu, real Pr, real mut, real Prt, const real A[3],
                                                         real out[5]) {
    const real mt = mu + mut;
    const real lam = real(-2.0 / 3.0) * mt * (gv[0][0] + gv[1][1] + gv[2][2]);
    real tA[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        real acc = lam * A[j];
#pragma unroll
        for (int i = 0; i < 3; ++i) acc += A[i] * (mt * (gv[i][j] + gv[j][i]));
        tA[j] = acc;
    }
    const real cp = real(kGamma * kRair / (kGamma - 1.0));
    const real k = mu * cp / Pr + mut * cp / Prt;
    out[0] = real(0);
    out[1] = tA[0];
    out[2] = tA[1];
    out[3] = tA[2];
    out[4] = q[1] * tA[0] + q[2] * tA[1] + q[3] * tA[2] - k * (A[0] * gT[0] + A[1] * gT[1] + A[2] * gT[2]);
}

// Per-wave walk over the cells of a persistent grid.  Workgroups are dispatched round-robin
// over the 8 XCDs (blockIdx % 8); each XCD walks one contiguous range of the cell list so
// that, with a locality-ordered list (schedule), the neighbour data of consecutive cells is
// shared in that XCD's L2.  The grid must be a multiple of 8 (or below 8): every list
// position is then visited exactly once.
struct CellWalk {
    int64_t it, hi, step;
    __device__ __forceinline__ CellWalk(int64_t n, int waves_per_block, int wid) {
        const int nxcd = gridDim.x >= 8 ? 8 : 1;
        const int xcd = blockIdx.x % nxcd, lb = blockIdx.x / nxcd;
        step = (int64_t)(gridDim.x / nxcd) * waves_per_block;
        it = n * xcd / nxcd + (int64_t)lb * waves_per_block + wid;
        hi = n * (xcd + 1) / nxcd;
    }
};

// cell of list position it: schedule[it] (a permutation of [0, n)) or it itself
__device__ __forceinline__ int64_t walk_cell(const int* schedule, int64_t it) {
    return schedule ? (int64_t)__builtin_amdgcn_readfirstlane(schedule[it]) : it;
}

// persistent grid of at most cap workgroups for n cells at wpb waves per workgroup, a
// multiple of 8 when at least 8
inline int walk_grid(int64_t n, int wpb, int cap) {
    int64_t need = (n + wpb - 1) / wpb;
    int64_t g = need < cap ? need : cap;
    g = g >= 8 ? g / 8 * 8 : g;
    return (int)(g > 0 ? g : 1);
}

}  // namespace afr
