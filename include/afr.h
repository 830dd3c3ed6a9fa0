/*
 * afr -- MI355X (gfx950) flux-reconstruction residual-and-update C ABI.
 *
 * Drop-in device backend for AutoFlowCFD's `--backend gpu` path.  All entry
 * points are plain C: device pointers, explicit sizes, an explicit
 * hipStream_t (passed as void*), int status (0 = ok) and afr_last_error().
 * Device buffers are owned by the caller (PyTorch-ROCm tensors on the Python
 * side); nothing here allocates device memory inside a hot call.
 *
 * Reference interfaces replaced (paths under the reference repository):
 *   afr_inviscid_residual  <- compute_inviscid_residual_fr_gpu
 *                             src/autoflowcfd/core/gpu/gpu_inviscid.py:31
 *                             (semantics of the CPU path
 *                             src/autoflowcfd/core/fr_residual/inviscid.py:139-345)
 *   afr_viscous_residual   <- compute_viscous_residual_fr_gpu
 *                             src/autoflowcfd/core/gpu/gpu_viscous.py:36
 *                             (semantics: src/autoflowcfd/core/fr_residual/viscous_flux.py:113-302)
 *   afr_primitives         <- conserved_to_primitive
 *                             src/autoflowcfd/core/fr_residual/inviscid.py:38-47
 *   afr_mesh_create        <- GPUArrayManager.upload_mesh_data + flat face arrays
 *                             src/autoflowcfd/core/gpu/array_manager.py:27-364,
 *                             src/autoflowcfd/core/gpu/gpu_solver_init.py:19-28
 */
#ifndef AFR_H
#define AFR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AFR_ABI_VERSION 6

enum { AFR_F32 = 0, AFR_F64 = 1 };

/* Device-resident description of one mesh at one order.  Every pointer is a
 * device pointer.  Cells are ordered prisms first ([0, n_prism)) then tets.
 * Real-valued arrays use the precision `dtype`.  Layouts (C = n_cells,
 * ns = (order+1)^3, nfp = (order+1)^2, nf = fine points, S = 5 slots):
 *   det        (C, ns)                 det(J) at solution points
 *   adj        (C, ns, 3, 3)           adj(J)[m][i] = det * J^{-1}[m][i]
 *   adj_fine   (C, nf, 3, 3)           adj(J) at over-integration points
 *   slot_nb    (C, S, 2) int32         source cells of each face slot (-1: none)
 *   slot_blk   (C, S, 2) int32         row-block ids into `rows`
 *   slot_mask  (C, S)    int32         bit fp set: flux point fp uses source 1; bit 30 set: the
 *                                       slot's source-0 rows have more than order+1 non-zeros
 *                                       (a scheduling hint of the viscous interface kernel:
 *                                       without it the wide rows are summed by one lane, with
 *                                       it in three parts; the two differ by round-off only)
 *   slot_bc    (C, S)    int32         BC code, -1 interior, -2 unused slot
 *   slot_fallback (C, S) int32         bit fp set: direction falls back to tn
 *   slot_tn    (C, S, 3)               role-signed true normal
 *   slot_adjrow (C, S, nfp, 3)         E @ adj[:, axis, :]
 *   rows       (n_rows, ns)            cross-cell interpolation rows
 *   rows_nnz   (n_rows) int32          non-zeros of each row (<= rows_knz)
 *   rows_idx   (n_rows, rows_knz) int32  source solution point of each non-zero
 *   rows_w     (n_rows, rows_knz)      weight of each non-zero
 *   c2f[t] (nf, ns), op[t] (ns, 3*nf), E[t] (S, nfp, ns), D[t] (3, ns, ns),
 *   filt[t] (ns, ns) for t = 0 prism, 1 tet
 *   gL, gR (order+1), fp_of_sp / coord_of_sp (3, ns) int32
 *   prism_schedule (n_prism) int32    processing order of the prisms (a permutation of
 *                                      [0, n_prism) in spatial locality order)
 *   prism_tab  packed prism factors (a,b)-triangle x c-line: c2f_ab, c2f_c,
 *              A_a = f2c_ab D_a, A_b = f2c_ab D_b, B_c = f2c_c D_c, f2c_c,
 *              f2c_ab, E_ab (3 quad slots), e_c (2 c slots), D_ab, D_c
 *   gcl        (C, ns, 3)              constant-state divergence sum_{m,j} Op[s][m,j] adj_fine[j][m][i]
 *   bc_params  (8, 5) FP64             per BC code reference state (always double);
 *                                      (n_cases, 8, 5) when cell_case is set
 *   cell_case  (C) int32, may be NULL  case index of every cell (batched multi-case meshes:
 *                                      disjoint cases with their own boundary states); the
 *                                      prisms of case i are [case_prism_off[i], case_prism_off[i+1]),
 *                                      its tets n_prism + [case_tet_off[i], case_tet_off[i+1]) and
 *                                      its wall records [case_wall_off[i], case_wall_off[i+1]).
 *                                      On such a mesh the pseudo-step entry points write one
 *                                      norm per case (norm: n_cases doubles).
 *   ref_state  (5) FP64, host values   primitive reference state Q_ref; the kernels carry
 *                                      dQ = Q - Q_ref in the storage precision and evaluate
 *                                      point fluxes / Riemann fluxes in FP64
 *   work_q     (C, ns, 5)              scratch for dQ = primitive - Q_ref
 *   work_grad  (C, ns, 12)             scratch for grad u (9, [i][j] = du_i/dx_j) and grad T (3)
 *   work_umean (C, 5)                  scratch for the per-cell mean |U| (outlier reference)
 *   work_rvol  (C, ns, 5)              scratch for the viscous volume term
 *   work_vpack (C, ns, 20)             scratch: per-point (dQ, grad u, grad T, mu_t, pad) pack
 *   dsp_w[t], dsp_i[t] (3, ns, W_t)    sparse rows of D[t] (W = 9 prism / 27 tet at P2; row m of
 *                                      direction m holds its non-zeros first, zero-padded)
 *   esp_w[t], esp_i[t] (5, nfp, K_t)   sparse rows of E[t] in slot order (K = nfp prism / ns tet)
 *   work_h     (C)                     scratch for the interior-penalty length mean(det)^(1/3)
 *   metric_flux_scale (C, ns)          sum_m |adj(J)[m, :]| (geometric time-step limit)
 *   volume     (C)                     cell volumes (GL quadrature)
 *   work_r, work_s1, work_s2 (C, ns, 5) pseudo-step scratch: residual, stage states
 *   work_dt    (C, ns)                 local time step
 *   work_part  (4096) FP64             partial sums of the residual norm
 * SST k-omega (optional, NULL / 0 when the mesh carries no turbulence model):
 *   n_records                          triangle face records (reference face order)
 *   rec_owner  (F) int32               owner cell of each record
 *   rec_oslot  (F) int32               owner face-slot index of the record's face
 *   rec_flags  (F) int32               bit 0: owner-primary interior record (has neighbour
 *                                      sources), bit 1: WALL boundary record
 *   rec_tn     (F, 3)                  record unit normal (owner -> neighbour / outward)
 *   cell_inc   (C, 8) int32            face-record incidences of each cell:
 *                                      record * 16 + role * 8 + cube-face code
 *                                      (role 0 owner, 1 neighbour), -1 unused
 *   work_pack  (C, 5, nfp, 8) owner face traces, work_nbp (C, ns, 4), work_tvol (C, ns, 2),
 *   work_tsrc  (C, ns, 3), work_crec (F, nfp, 2)
 * Wall forces (optional):
 *   n_wall_rec                         owner-primary WALL face records
 *   wrec_owner, wrec_oslot (W) int32   owner cell and owner face slot
 *   wrec_geom  (W, 4) FP64             record unit normal, record area
 *   w1d        (order+1) FP64          Gauss-Legendre weights
 *   work_fpart (W, 6) FP64             per-record force partial sums
 * State storage:
 *   state_ref  (n_cases, 5) FP64 or NULL  conserved offset of each case: every state array of
 *                                      this mesh (U, U0, the stage states) then holds
 *                                      U - state_ref[case of the cell], and every kernel adds the
 *                                      offset back in FP64 when it loads a point.  NULL: the
 *                                      arrays hold U itself.  With the case's free stream as the
 *                                      offset, FP32 storage resolves the flow's variation about
 *                                      the free stream instead of rounding the dimensional state
 *                                      (rho E ~ 2.5e5 rounds to 0.016).  Without cell_case the
 *                                      one row is copied at afr_mesh_create (the kernels take it
 *                                      by value); with cell_case the table is read by the kernels
 *                                      and must stay valid and unchanged while the mesh exists.
 *                                      afr_mesh_create also reads cell_case once: when it ascends
 *                                      over the prisms and over the tets (the concatenated
 *                                      layout), a wavefront whose cells share one case reads that
 *                                      case's row once; any other cell_case gives the same
 *                                      results through per-cell lookups.
 */
typedef struct AfrMeshDesc {
    int32_t abi_version;
    int32_t order;
    int32_t dtype;
    int32_t n_rows;
    int64_t n_cells;
    int64_t n_prism;
    const void *det, *adj, *adj_fine;
    const int32_t *slot_nb, *slot_blk, *slot_mask, *slot_bc, *slot_fallback;
    const void *slot_tn, *slot_adjrow;
    const void *rows;
    const int32_t *rows_nnz, *rows_idx;
    const void *rows_w;
    int32_t rows_knz;
    int32_t pad_;
    const void *c2f[2], *op[2], *E[2], *D[2], *filt[2];
    const void *gL, *gR;
    const int32_t *fp_of_sp, *coord_of_sp;
    const void *prism_tab;
    const int32_t *prism_schedule;
    const void *gcl;
    const double *bc_params;
    const int32_t *cell_case;
    int32_t n_cases;
    int32_t pad2_;
    const int64_t *case_prism_off; /* HOST (n_cases + 1), copied at afr_mesh_create */
    const int64_t *case_tet_off;   /* HOST (n_cases + 1) */
    const int64_t *case_wall_off;  /* HOST (n_cases + 1) over the wall records; NULL without walls */
    double ref_state[5];
    void *work_q;
    void *work_grad;
    void *work_umean;
    void *work_rvol;
    void *work_h;
    void *work_vpack;
    const void *dsp_w[2], *esp_w[2];
    const int32_t *dsp_i[2], *esp_i[2];
    const void *metric_flux_scale, *volume;
    void *work_r, *work_s1, *work_s2, *work_dt;
    double *work_part;
    int64_t n_records;
    const int32_t *rec_owner, *rec_oslot, *rec_flags;
    const void *rec_tn;
    const int32_t *cell_inc;
    void *work_pack, *work_nbp, *work_tvol, *work_tsrc, *work_crec;
    int64_t n_wall_rec;
    const int32_t *wrec_owner, *wrec_oslot;
    const double *wrec_geom, *w1d;
    double *work_fpart;
    /* persistent-grid sizes (workgroups) per kernel, indexed by AfrKernel; 0 = built-in default.
     * Set by afr_mesh_set_grid (the autotuner), not read from the caller's descriptor. */
    int32_t grid_cap[8];
    /* kernel variants (tile shapes) per kernel, indexed by AfrKernel; 0 = the built-in default.
     * Set by afr_mesh_set_variant (the autotuner), not read from the caller's descriptor. */
    int32_t variant[8];
    const double *state_ref;
} AfrMeshDesc;

/* Kernels whose persistent grid size is tunable (afr_mesh_set_grid). */
enum AfrKernel {
    AFR_K_PRISM_INVISCID = 0,
    AFR_K_TET_INVISCID = 1,
    AFR_K_VISC_VOLUME = 2,
    AFR_K_VISC_IFACE = 3,
    AFR_K_SST_CELL = 4,
    AFR_K_SST_FACE = 5,
    AFR_K_SST_UPDATE = 6
};

typedef struct AfrMesh AfrMesh;

/* Inputs/outputs of the FP64 metric precompute (afr_metrics).  nodes (N,3)
 * FP64, prism_conn (n_prism,6) / tet_conn (n_tet,4) int32 (orientation-fixed),
 * sps (ns,3) / fine (nf,3) / w_sps (ns) FP64 reference points and weights,
 * op64[t] (ns, 3*nf) and E64[t] (5, nfp, ns) FP64 tables in slot order,
 * slot_tn64 (C,5,3) FP64.  Outputs in the storage precision `dtype`: det,
 * adj, adj_fine, gcl, metric_flux_scale, slot_adjrow (layouts as in
 * AfrMeshDesc); FP64 volume (C) and det_mean (C); int32 slot_fallback (C,5);
 * bad_det (1 int32, set non-zero if any det(J) <= 0; caller zeroes it);
 * sps_xyz (C, ns, 3) FP64 solution-point positions, optional (NULL: not written). */
typedef struct AfrGeomDesc {
    int32_t abi_version;
    int32_t order;
    int32_t dtype;
    int32_t pad_;
    int64_t n_cells;
    int64_t n_prism;
    const double *nodes;
    const int32_t *prism_conn, *tet_conn;
    const double *sps, *fine, *w_sps;
    const double *op64[2], *E64[2];
    const double *slot_tn64;
    void *det, *adj, *adj_fine, *gcl, *metric_flux_scale, *slot_adjrow;
    double *volume, *det_mean;
    int32_t *slot_fallback;
    int32_t *bad_det;
    double *sps_xyz;
} AfrGeomDesc;

/* Library ABI version (AFR_ABI_VERSION). */
int afr_version(void);

/* Copies the message of the last failed call on this thread into buf. */
int afr_last_error(char *buf, size_t n);

/* Validates and records a mesh description (no device allocation). */
int afr_mesh_create(const AfrMeshDesc *desc, AfrMesh **out);
int afr_mesh_destroy(AfrMesh *mesh);

/* FP64 metric precompute on the device (replaces the per-cell Python loops of
 * src/autoflowcfd/grid/high_order/high_order_mesh_order.py:47-99,152-160 and
 * the flat-face adjugate extrapolation of inviscid_kernel.py:131-133). */
int afr_metrics(const AfrGeomDesc *geom, void *stream);

/* dQ = conserved_to_primitive(U) - ref (ref: 5 host doubles, may be NULL) for n_points points. */
int afr_primitives(const void *U, void *Q, int64_t n_points, int32_t dtype, const double *ref, void *stream);

/* Per-step state preparation of one mesh: dQ = primitive(U) - ref_state
 * (C, ns, 5) and umean[c][v] = mean over solution points of |U[c][s][v]| (C, 5),
 * the reference magnitude of the outlier test (troubled_cell.py:371-380). */
int afr_prepare_state(const AfrMesh *mesh, const void *U, void *dQ, void *umean, void *stream);

/* Inviscid FR residual dU/dt (C, ns, 5): over-integrated volume term, AUSM+up
 * interface correction, optional outlier suppression.  U is (C, ns, 5).
 * Uses the mesh work buffers work_q and work_umean. */
int afr_inviscid_residual(const AfrMesh *mesh, const void *U, void *R, int32_t suppress, void *stream);

/* Same as afr_inviscid_residual from the prepared state (dQ, umean) of
 * afr_prepare_state. */
int afr_inviscid_from_primitives(const AfrMesh *mesh, const void *dQ, const void *umean, void *R,
                                 int32_t suppress, void *stream);

/* Viscous BR1 FR residual dU/dt (C, ns, 5); mu_t (C, ns) or NULL.
 * mu, Pr, Pr_t as in the reference (1.8e-5, 0.72, 0.9 by default).  Uses the
 * work buffers work_q, work_umean, work_grad, work_rvol and work_h; on return
 * work_q / work_umean hold the prepared state of U and work_grad its gradients. */
int afr_viscous_residual(const AfrMesh *mesh, const void *U, const void *mu_t, void *R, double mu,
                         double Pr, double Pr_t, int32_t suppress, void *stream);

/* The two phases of afr_viscous_residual, for callers that interleave other work:
 * afr_viscous_volume writes the prepared state (work_q, work_umean), the gradients
 * (work_grad), the viscous volume term (work_rvol) and work_h from U;
 * afr_viscous_interface adds the BR1 interface correction to work_rvol, applies the
 * outlier zeroing when suppress != 0 and writes (accumulate == 0) or adds
 * (accumulate != 0) the result to R. */
int afr_viscous_volume(const AfrMesh *mesh, const void *U, const void *mu_t, double mu, double Pr, double Pr_t,
                       void *stream);
int afr_viscous_interface(const AfrMesh *mesh, const void *U, const void *mu_t, void *R, double mu, double Pr,
                          double Pr_t, int32_t suppress, int32_t accumulate, void *stream);

/* Full Navier-Stokes FR residual dU/dt = inviscid + viscous (each with its own
 * outlier zeroing when suppress != 0), as the reference solver sums them
 * (src/autoflowcfd/core/fr_solver/solver.py:360-420).  mu_t may be NULL. */
int afr_residual(const AfrMesh *mesh, const void *U, const void *mu_t, void *R, double mu, double Pr,
                 double Pr_t, int32_t suppress, void *stream);

/* Local time step dt (C, ns) of the conserved state U (C, ns, 5):
 * min(advective, viscous, geometric) limits with CFL 0.1 and the order factors
 * 1/(2P+1), 1/(2P+1)^2 (src/autoflowcfd/core/fr_solver/cfl.py:12-147); mu_t may be NULL. */
int afr_local_time_step(const AfrMesh *mesh, const void *U, const void *mu_t, void *dt, double mu, void *stream);

/* As afr_local_time_step with the eddy viscosity given as nu_t: mu_t = rho(U) nu_t. */
int afr_local_time_step_nut(const AfrMesh *mesh, const void *U, const void *nu_t, void *dt, double mu, void *stream);

/* One Shu-Osher stage U_out = c0 U0 + c1 Ua + c2 dt R followed by the positivity
 * projection (p_floor 1) and the modal filter of the first 5 variables
 * (src/autoflowcfd/core/time_integration/base.py:250-335). U_out may alias U0 or Ua. */
int afr_rk_stage(const AfrMesh *mesh, const void *U0, const void *Ua, const void *R, const void *dt, void *U_out,
                 double c0, double c1, double c2, void *stream);

/* RMS norm sqrt(sum R^2 / (C ns n_vars_total)) accumulated in FP64 into out[0]
 * (a device pointer; src/autoflowcfd/core/fr_solver/state.py:99-107).  n_vars_total
 * counts the turbulence components of the reference state (7 with SST), whose
 * residual entries are zero. */
int afr_residual_norm_f64(const AfrMesh *mesh, const void *R, int32_t n_vars_total, double *out, void *stream);

/* afr_residual_norm_f64 without the mesh's work_part buffer: the FP64 partial sums go to a
 * file-scope device array of the library (4096 doubles = 32 KiB per GPU, held until the process
 * ends), so the call allocates nothing and needs no workspace.  No counterpart in the reference;
 * it is for callers that cannot allocate or bring a workspace (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same partition and summation order as
 * afr_residual_norm_f64, hence the same bits.  Every length of R fits: the partial count is
 * capped at the array's 4096 entries whatever the mesh size, and each call writes every entry
 * it reads, so nothing needs clearing.  Calls that overlap on one GPU (two streams, two threads)
 * race on that array: serialise them.  Errors: null arguments, n_vars_total < 5, no current HIP
 * device, the device address of the array not available. */
int afr_residual_norm_f64_noalloc(const AfrMesh *mesh, const void *R, int32_t n_vars_total, double *out,
                                  void *stream);

/* One steady SSP-RK3 pseudo-step of the Navier-Stokes state U (C, ns, 5), in place:
 * dt_local, R0 = residual(U), three stages with positivity and filter, and the RMS
 * norm of R0 into norm[0] (device pointer; on a batched mesh norm[i] per case, as
 * afr_residual_norm_cases), as src/autoflowcfd/core/fr_solver/step.py:15-221.
 * mu_t (C, ns) is frozen over the step (may be NULL); mu == 0 with mu_t == NULL skips
 * the viscous terms.  Uses work_r, work_s1, work_s2, work_dt, work_part. */
int afr_ssp_rk3_step(const AfrMesh *mesh, void *U, const void *mu_t, double mu, double Pr, double Pr_t,
                     int32_t n_vars_total, double *norm, void *stream);

/* afr_ssp_rk3_step with one time step dt for every solution point instead of dt_local (the
 * global-dt stepping of the reference's unsteady validation runs,
 * tests/validation/test_tgv.py:70-92 with src/autoflowcfd/core/fr_solver/step.py:15-221). */
int afr_ssp_rk3_step_dt(const AfrMesh *mesh, void *U, const void *mu_t, double dt, double mu, double Pr,
                        double Pr_t, int32_t n_vars_total, double *norm, void *stream);

/* Distance from each of n_pts FP64 points (n_pts, 3) to the nearest of n_wall FP64
 * wall-node positions (n_wall, 3), written in precision dtype to d_out (n_pts)
 * (src/autoflowcfd/core/fr_solver/turbulence.py:112-121: KD-tree to wall nodes). */
int afr_wall_distance(const double *sp_xyz, int64_t n_pts, const double *wall_xyz, int32_t n_wall, void *d_out,
                      int32_t dtype, void *stream);

/* SST k-omega step of the turbulence fields k, omega (C, ns) in place with the
 * local time step dt (C, ns), evaluated at the conserved state U: sources with
 * F1/F2 blending and eddy viscosity, record-based transport, point-implicit update
 * and floors (src/autoflowcfd/core/fr_solver/turbulence.py:178-330,
 * src/autoflowcfd/core/turbulence/{sst,transport,transport_kernel}.py).  Writes
 * nu_t (C, ns) and mu_t = rho(U) nu_t (C, ns), the eddy viscosity of the step.  With
 * rates != NULL the fields are left unchanged and rates (C, ns, 2) receives dk/dt and
 * domega/dt of the update (the turbulence part of a residual evaluation). */
int afr_sst_step(const AfrMesh *mesh, const void *U, void *k, void *omega, const void *d_wall, const void *dt,
                 void *nu_t, void *mu_t, void *rates, double mu_molecular, void *stream);

/* One steady pseudo-step with the SST model, as src/autoflowcfd/core/fr_solver/step.py:15-221:
 * dt_local from U with mu_t = rho(U) nu_t of the previous step (nu_t in/out, zero
 * initially), the SST step (k, omega, nu_t, mu_t updated), then the SSP-RK3 stages of U
 * with mu_t frozen; norm[0] (device) = RMS of R(U) over C ns n_vars_total entries (on a batched
 * mesh norm[i] per case). */
int afr_sst_rk3_step(const AfrMesh *mesh, void *U, void *k, void *omega, const void *d_wall, void *nu_t, void *mu_t,
                     double mu, double Pr, double Pr_t, int32_t n_vars_total, double *norm, void *stream);

/* Residual outlier zeroing of a field R (n_cells, ns, nv) in place, as
 * suppress_residual_outliers (src/autoflowcfd/core/fr_operators/troubled_cell.py:382-408) with its
 * default factor 1e4 and field floor 1e-9: R[c][s][v] = 0 where |R| > 1e4 max(median_s |R[c][:][v]|,
 * 1e-9 mean_s |U[c][:][v]|, 1e-300); the median as troubled_cell.py:347-379 takes it (sorted, NaN
 * last, mean of the two middle values for even ns), a NaN median zeroes nothing.  U (n_cells, ns,
 * u_stride) is the reference field, its first nv variables used.  ns 1, 8, 27 or 64; nv 1, 2, 5
 * or 7; dtype AFR_F32 / AFR_F64 for both arrays.  The residual entry points apply the same rule
 * inside their kernels when `suppress` is set. */
int afr_suppress_residual_outliers(void *R, const void *U, int64_t n_cells, int32_t ns, int32_t nv,
                                   int32_t u_stride, int32_t dtype, void *stream);

/* Persistent grid size of one kernel on this mesh: grid_cap workgroups (0 = the built-in
 * default; otherwise 8 .. 65536, rounded down to a multiple of the 8 XCDs).  Results do not
 * depend on it; the autotuner (autoflowcfd_amd/autotune.py) picks it by timing.  No
 * reference counterpart (the reference launches CuPy kernels with fixed grids). */
int afr_mesh_set_grid(AfrMesh *mesh, int32_t kernel, int32_t grid_cap);
int afr_mesh_get_grid(const AfrMesh *mesh, int32_t kernel);

/* Tile shape of one kernel on this mesh (no reference counterpart; picked by the autotuner with
 * the grid size).  AFR_K_PRISM_INVISCID: 0 = the own cell's state staged in LDS, the slot sources
 * read from global memory; 1 = the own cell and its five slot sources staged in LDS (more LDS per
 * wavefront, fewer waves per SIMD).  AFR_K_SST_CELL: 0 = two prisms per wavefront (half-waves of
 * 32 lanes, P <= 2 prisms), 1 = one cell per wavefront.  Other kernels: 0 only.  Every variant
 * computes the same values bit for bit. */
int afr_mesh_set_variant(AfrMesh *mesh, int32_t kernel, int32_t variant);
int afr_mesh_get_variant(const AfrMesh *mesh, int32_t kernel);

/* Per-case RMS norms of a batched mesh (cell_case set): out[i] = sqrt(sum over the cells of
 * case i of R^2 / (cells_i ns n_vars_total)); out: n_cases doubles (device).  Each case is
 * reduced in the same order as afr_residual_norm_f64 over that case alone: bit-identical values.
 * Per-case form of src/autoflowcfd/core/fr_solver/state.py:99-107. */
int afr_residual_norm_cases(const AfrMesh *mesh, const void *R, int32_t n_vars_total, double *out, void *stream);

/* Per-case wall forces of a batched mesh: out (n_cases, 6) doubles (device), bit-identical to
 * afr_forces_f64 of each case alone.  Per-case form of
 * src/autoflowcfd/postprocess/fr_coefficients.py:108-112. */
int afr_forces_cases_f64(const AfrMesh *mesh, const void *U, const void *mu_t, double mu, double *out,
                         void *stream);

/* FP64 force on the WALL boundary of the state U (C, ns, 5) with eddy viscosity mu_t
 * (C, ns, may be NULL): out[0..2] pressure force sum p n w, out[3..5] viscous force
 * -sum (tau . n) w (device pointer, 6 doubles), as
 * src/autoflowcfd/postprocess/fr_coefficients.py:27-152.  Cd = F_x/(q A), Cl = F_z/(q A). */
int afr_forces_f64(const AfrMesh *mesh, const void *U, const void *mu_t, double mu, double *out, void *stream);

/* The P0 phase of the order continuation (src/autoflowcfd/core/utils/order_continuation.py:207-425)
 * on one mesh: plain arrays, no library object.  Per cell: det J, adj J (3x3) and the metric flux
 * scale at the P0 point (the centre of the reference cube), vol1 = the one-point volume 8 det J
 * (the local time step's h, cfl.py:96-101) in the storage precision, and cell_volume (FP64) = the
 * target-order volume of the finite-volume residual.  Per face record: owner, neighbour (-1 on
 * the boundary), bc = type | (row << 4) on the boundary (-1 inside), flags (bit 0 owner-primary,
 * bits 1-2 the owner's face axis, bit 3 its side +1), geom (FP64: unit normal owner -> neighbour,
 * area).  cell_inc (C, 8): record * 16 + role * 8 + face code (role 1: the cell is the
 * neighbour), -1 padded; cell_ghost_rec (C): the cell's last boundary record (-1 none).
 * bc_params (8, 5) FP64.  Work buffers (C, 5) x 3, dt (C), work_part 4096 doubles. */
typedef struct AfrP0Desc {
    int32_t abi_version;
    int32_t dtype;
    int64_t n_cells;
    int64_t n_records;
    const void *det, *adj, *mfs, *vol1;
    const double *cell_volume;
    const int32_t *rec_owner, *rec_nb, *rec_bc, *rec_flags;
    const double *rec_geom;
    const int32_t *cell_inc, *cell_ghost_rec;
    const double *bc_params;
    void *work_r, *work_s1, *work_s2, *work_dt;
    double *work_part;
    /* SST (afr_p0_sst_step only, else NULL): per record the neighbour sources of its owner side
     * at the one flux point (rec_src: two cells, -1 none; rec_w: their FP64 weights, 0 or 1: a prism
     * quad's point lies on its diagonal and goes to the half the reference assigns it); flags bit 4
     * marks WALL records; d_wall (C) the wall distance; work_sst (C, 8) FP64 scratch. */
    const int32_t *rec_src;
    const double *rec_w;
    const void *d_wall;
    double *work_sst;
    /* (5) FP64 or NULL: the conserved offset of the stored state (as AfrMeshDesc.state_ref, one
     * case) */
    const double *state_ref;
} AfrP0Desc;

/* P0 residual of the conserved state U (C, 1, 5) into R: the finite-volume inviscid residual of
 * src/autoflowcfd/core/fr_residual/inviscid_p0.py:17-143 (each boundary record with the ghost
 * state its owner cell keeps, that of the cell's last boundary record, inviscid_p0.py:118-143)
 * plus the one-point viscous residual (viscous_p0_kernel.py:26-285: the boundary interior penalty,
 * the only term a constant state leaves).  mu_t (C) may be NULL. */
int afr_p0_residual(const AfrP0Desc *desc, const void *U, const void *mu_t, double mu, void *R, void *stream);

/* One P0 SSP-RK3 pseudo-step of U (C, 1, 5) in place (dt_local at P0, three stages with
 * positivity; the P0 filter is the identity) and the RMS norm of R0 into norm[0] (device). */
int afr_p0_ssp_rk3_step(const AfrP0Desc *desc, void *U, const void *mu_t, double mu, int32_t n_vars_total,
                        double *norm, void *stream);

/* One P0 pseudo-step with the SST model (src/autoflowcfd/core/fr_solver/step.py:15-221 on the P0
 * phase of the order continuation): the local time step with the previous step's eddy viscosity
 * (nu_t in, mu_eff = mu + rho nu_t); the SST turbulence step with the one-point operators
 * (zero gradients: sources from destruction only, nu_t = a1 k / max(a1 omega, F2 1e-10) with the
 * viscosity-ratio cap; transport from the face records alone: upwind convection and the BR1-type
 * diffusion correction distributed with g' = -1/2, +1/2 over det J, Neumann ghost on records without
 * neighbour sources, Dirichlet-zero k on WALL records; point-implicit update with floors),
 * src/autoflowcfd/core/turbulence/sst.py:203-479 and transport.py:122-520 at P0; then the
 * SSP-RK3 stages of U with mu_t = rho nu_t frozen.  k, omega, nu_t, mu_t (C) updated in place;
 * norm[0] = RMS of R(U) over C n_vars_total entries. */
int afr_p0_sst_step(const AfrP0Desc *desc, void *U, void *k, void *omega, void *nu_t, void *mu_t, double mu,
                    int32_t n_vars_total, double *norm, void *stream);

/* X_new (C, ns_new, n_vars) = W (ns_new, ns_old, FP64, device) applied to X_old (C, ns_old, n_vars)
 * per cell: the interpolation between orders of order_continuation.py:14-57.  ns <= 64. */
int afr_interpolate_order(const void *X_old, void *X_new, const double *W, int64_t n_cells, int32_t ns_old,
                          int32_t ns_new, int32_t n_vars, int32_t dtype, void *stream);

/* Face topology of a prism / tet mesh on the GPU, as FaceExtractor.extract_faces_mixed
 * (src/autoflowcfd/grid/mesh_gen/extraction/face_extractor.py:296-376) with the cube-face codes of
 * src/autoflowcfd/grid/connectivity/face_connectivity.py:53-231: every prism contributes 8
 * triangles (its quads split on the global-node-id diagonal rule, face_extraction_kernels.py:247-322),
 * every tet 4; the records are the distinct sorted node triples in ascending order, owner = the
 * first triangle in cell order, neighbour = the second (-1 on the boundary); normal (unit, away from
 * the owner's vertex centroid), area and centre in FP64 over the sorted triple
 * (face_geometry_finalize.py:92-160).  Inputs: nodes (n_nodes, 3) FP64, prism_conn (n_prism, 6),
 * tet_conn (n_tet, 4) int32 node ids.  Outputs, each with room for 8 n_prism + 4 n_tet records:
 * owner, neighbor, owner_code, neighbor_code (int32), node_ids (3 per record, sorted), normal (3),
 * area, center (3).  n_faces[0] (device) = the record count; status[0] (device) = 0, or bit 0 set
 * when a triple is shared by more than two cells, bit 1 when a side matches no cube face or
 * several.  work: afr_face_topology_workspace bytes of device scratch. */
typedef struct AfrTopoDesc {
    int32_t abi_version;
    int32_t pad_;
    int64_t n_nodes, n_prism, n_tet;
    const double *nodes;
    const int32_t *prism_conn, *tet_conn;
    void *work;
    size_t work_bytes;
    int32_t *owner, *neighbor, *owner_code, *neighbor_code, *node_ids;
    double *normal, *area, *center;
    int32_t *n_faces, *status;
} AfrTopoDesc;

int afr_face_topology_workspace(int64_t n_prism, int64_t n_tet, size_t *bytes);
int afr_face_topology(const AfrTopoDesc *desc, void *stream);

/* Cell-centric face slots on the GPU: the tables of the slot form of the reference's flat face
 * geometry (src/autoflowcfd/fr/face_flux_points_merge.py:83-124, the primary side of every cube
 * face; face_flux_points.py:164-236, the cross-interpolation rows keyed by the face correspondence;
 * face_flux_points_data.py:145-217, prism quads split between two cells).  Inputs: the mesh
 * (nodes FP64, prism_conn, tet_conn int32) and its face records (owner, neighbor, owner_code,
 * neighbor_code, bc, bc_row (may be NULL) int32; normal, translation (3 per record) FP64).  Outputs
 * per slot (C = n_prism + n_tet cells x 5 slots): nb (2: source cells, -1 none), blk (2: row block
 * of a keyed slot's first source, 0 otherwise; split slots are completed by the caller), slot_bc
 * (BC code, -1 interior, -2 unused), slot_bc_row (may be NULL), tn (3, role-signed true normal),
 * trans (3, role-signed periodic translation), other (4: other cell 0, its face code, other cell 1,
 * its face code), rec (the record), slot_flags (bit 0 keyed, bit 1 split); per distinct key (row
 * block): rep (its first slot) and run_start (its first position in sorted_slot); sorted_slot
 * (every slot, keyed slots by ascending key first); counts[0] = distinct keys, counts[1] = keyed
 * slots; status[0] = 0, bit 0 when a real cube face is not the primary side of exactly one record,
 * bit 1 when the corners of a periodic pair do not coincide.  work: afr_face_slots_workspace bytes. */
typedef struct AfrSlotDesc {
    int32_t abi_version;
    int32_t pad_;
    int64_t n_prism, n_tet, n_records;
    const double *nodes;
    const int32_t *prism_conn, *tet_conn;
    const int32_t *owner, *neighbor, *owner_code, *neighbor_code, *bc, *bc_row;
    const double *normal, *translation;
    void *work;
    size_t work_bytes;
    int32_t *nb, *blk, *slot_bc, *slot_bc_row;
    double *tn, *trans;
    int32_t *other, *rec, *slot_flags, *rep, *run_start, *sorted_slot, *counts, *status;
} AfrSlotDesc;

int afr_face_slots_workspace(int64_t n_cells, size_t *bytes);
int afr_face_slots(const AfrSlotDesc *desc, void *stream);

/* HIP graph of afr_* calls (no reference counterpart: the reference launches every kernel of a
 * pseudo-step from Python, src/autoflowcfd/core/gpu/gpu_solver.py:step).  afr_graph_begin starts
 * capturing the work this host thread issues on `stream` (thread-local capture: other threads
 * may keep launching); afr_graph_end stops and instantiates it; afr_graph_launch replays it on a
 * stream with the pointers and scalars it was captured with, e.g. one afr_ssp_rk3_step or
 * afr_sst_rk3_step per launch, one launch per pseudo-step.  The afr_* entry points allocate
 * nothing, issue no memset or copy and do not synchronise, so each of them can be captured.
 * `stream` must be a created stream, not the null stream. */
typedef struct AfrGraph AfrGraph;
int afr_graph_begin(void *stream);
int afr_graph_end(void *stream, AfrGraph **out);
int afr_graph_launch(AfrGraph *graph, void *stream);
int afr_graph_destroy(AfrGraph *graph);

#ifdef __cplusplus
}
#endif
#endif /* AFR_H */
