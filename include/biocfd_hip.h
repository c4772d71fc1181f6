/*
 * biocfd_hip.h -- C ABI of the MI355X (gfx950) implementation of the Bio-CFD
 * per-time-step hot path.
 *
 * Every entry point returns 0 on success and a negative code on failure; the
 * message of the last failure is available from biocfd_last_error().  Blocks
 * are identified by the reference's 1-based block number; arrays crossing the
 * boundary use the reference's host layout (Fortran column-major, 1-based
 * with one ghost layer: field(i,j,k), i = 1..nx+2, at
 * f[(i-1) + (nx+2)*((j-1) + (ny+2)*(k-1))]; index lists (count,3) int64
 * column-major; tag arrays int64).  Device copies stay resident in HBM between
 * calls; uploads/downloads are explicit.
 *
 * Each compute entry point names the reference procedure it replaces
 * (reference file:line).
 */
#ifndef BIOCFD_HIP_H
#define BIOCFD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- identifiers ------------------------------------------------------- */
enum biocfd_field { BF_U = 0, BF_V, BF_W, BF_UT, BF_VT, BF_WT, BF_P, BF_B, BF_PC, BF_PCO, BF_NFIELD };
enum biocfd_tag { BT_CELL = 0, BT_CELL2, BT_NODEIDTAG, BT_CELL_N, BT_CELL_PR, BT_NTAG };
/* grid arrays in Block_t order (Blocks.f90:26-27) */
enum biocfd_grid { BG_X1 = 0, BG_Y1, BG_Z1, BG_DX, BG_DY, BG_DZ, BG_XU, BG_YU, BG_ZU, BG_XV, BG_YV,
                   BG_ZV, BG_XW, BG_YW, BG_ZW, BG_XP, BG_YP, BG_ZP, BG_NGRID };
enum biocfd_list { BL_FLUID = 0, BL_RED, BL_BLACK, BL_SOLID, BL_INTERCEPTED, BL_TS, BL_NLIST };
/* per-IB-cell arrays: nearest element indices (Blocks.f90:49) */
enum biocfd_nel { BN_P = 0, BN_U1, BN_U2, BN_V1, BN_V2, BN_W1, BN_W2, BN_NNEL };
/* per-element arrays (Blocks.f90:50-52) */
enum biocfd_elem { BE_XCENT = 0, BE_YCENT, BE_ZCENT, BE_COSA, BE_COSB, BE_COSG, BE_ELEN, BE_NELEM };
/* TS ghost arrays (Blocks.f90:55-58) */
enum biocfd_ghost { BGH_P = 0, BGH_PT, BGH_U1, BGH_U1T, BGH_U2, BGH_U2T, BGH_V1, BGH_V1T, BGH_V2,
                    BGH_V2T, BGH_W1, BGH_W1T, BGH_W2, BGH_W2T, BGH_N };
/* counts (Blocks.f90:42-43) */
enum biocfd_count { BC_FLUID = 0, BC_SOLID, BC_IB, BC_RED, BC_BLACK, BC_TS, BC_N };
/* body kinematic scalars read by the forcing routines (Blocks.f90:14-15, 72-73) */
enum biocfd_kin { BK_XDOT = 0, BK_YDOT, BK_YDDOT, BK_THETADOT1, BK_THETADDOT1, BK_THETADOT2,
                  BK_THETADDOT2, BK_ALPHADOT, BK_PIVX, BK_PIVY, BK_PIVZ, BK_N };
/* 3rd-order upwind tables, Blocks.f90:29-31 order */
enum biocfd_table { BTB_UU = 0, BTB_VV, BTB_WW, BTB_UV, BTB_UW, BTB_VU, BTB_VW, BTB_WU, BTB_WV, BTB_N };

/* ---- runtime ----------------------------------------------------------- */
const char *biocfd_last_error(void);
/* Select the HIP device and create the runtime's stream (no reference
 * counterpart: replaces set_gpu, gpu_helpers.F90:17-40, on one GPU). */
int biocfd_init(int device);
int biocfd_finalize(void);
int biocfd_synchronize(void);

/* Time the red-black SOR sweeps with HIP events on the runtime's stream (no
 * reference counterpart: measurement hook).  read: total GPU milliseconds of
 * the SOR iteration loops since enable, iterations executed, and interior
 * cells x iterations swept. */
int biocfd_timing_enable(int on);
int biocfd_timing_read(double *sor_ms, int64_t *sor_iters, int64_t *sor_cell_iters);
/* Device-to-device copy of `bytes` bytes, `reps` times (no reference
 * counterpart: calibrates HBM byte counters against a copy of known size). */
int biocfd_copy_probe(int64_t bytes, int reps);

/* ---- block data (Block_t, Blocks.f90:13-81; AllocateArrays.f90:15-57) ---- */
int biocfd_block_create(int blk, int64_t nx, int64_t ny, int64_t nz, double omega, int64_t cintp,
                        double dx, double dy, double dz);
int biocfd_block_destroy(int blk);
int biocfd_grid_upload(int blk, int which, const double *host, int64_t n);
int biocfd_field_upload(int blk, int field, const double *host);
int biocfd_field_download(int blk, int field, double *host);
int biocfd_tag_upload(int blk, int tag, const int64_t *host);
int biocfd_tag_download(int blk, int tag, int64_t *host);
int biocfd_counts(int blk, int64_t out[BC_N]);
/* materialise an index list in the reference's (count,3) int64 layout and
 * serial i-fastest order (cellCount_solid, Search.F90:653-741) */
int biocfd_list_download(int blk, int which, int64_t *host, int64_t count);
int biocfd_nel_download(int blk, int which, int64_t *host, int64_t count);
int biocfd_normdis_download(int blk, int which, double *host, int64_t count);
int biocfd_ghost_download(int blk, int which, double *host, int64_t count);
int biocfd_elem_download(int blk, int which, double *host, int64_t count);
int biocfd_index_ts_download(int blk, int64_t *host, int64_t count);
int biocfd_coef_download(int blk, int axis, double *host, int64_t n);       /* Acx/Acy/Acz (n,3) */
int biocfd_table_download(int blk, int which, int is_ck, double *host, int64_t ncols);
/* SOR state of the last poissonSolver call: nIterPcor, derr2, derrStdSt */
int biocfd_sor_state(int blk, int64_t *n_iter, double *derr2, double *derr_std_st);

/* ---- failure detection (no reference counterpart: the reference checks
 * nothing; an SOR call that reaches pcItaMax ends silently, PcorVcor.F90:181,
 * 351-375).  All off by default, so the reference's behaviour is kept. ---- */
/* Red-black SOR calls of the last poissonSolver call on this block that ended
 * on their iteration budget (nIterPcor > pcItaMax) rather than their
 * tolerance (counted on the device by the end-of-call kernels; a block is
 * solved up to twice per poissonSolver call). */
int biocfd_sor_exhausted(int blk, int *count);
/* Scan u, v, w, p of a block (every cell, ghost layer included): the number
 * of non-finite values, the largest |u|, |v|, |w| among the finite ones, and
 * where[0..3] = field (0 u, 1 v, 2 w, 3 p), i, j, k (1-based) of the first
 * non-finite value in (field, k, j, i) order, or all -1. */
int biocfd_health_check(int blk, int64_t *n_nonfinite, double *vel_max, int64_t where[4]);
/* check_after_solve != 0: biocfd_poisson_solver ends with the scan of its
 * blocks and returns -3 if a value is non-finite or a velocity component
 * exceeds vel_limit (vel_limit 0: no limit), and, with exhausted_is_error,
 * -4 if an SOR call reached pcItaMax; the error text names the block (and
 * the first non-finite cell).  The fields are left as the solve wrote them.
 * Cost: one read of four fields per block and one stream synchronisation. */
int biocfd_health_config(int check_after_solve, double vel_limit, int exhausted_is_error);
/* Test hook: the nth device allocation from now (nth >= 1) fails as a failed
 * hipMalloc would, with its error text marked as injected; 0 disarms it.  An
 * entry point whose allocation fails returns -1 and leaves the library usable
 * (a block whose creation fails is not created). */
int biocfd_inject_alloc_failure(int nth);

/* ---- immersed body (readSurfaceMeshGmsh target arrays, ReadInput.f90:375-420) ---- */
int biocfd_body_create(int blk, int64_t n_nodes, int64_t n_elems, const int64_t *el_p1,
                       const int64_t *el_p2, const int64_t *el_p3, const int64_t *surf_id,
                       const int64_t search_box[6]);

/* ---- set-up numerics (host-side, results kept on the device) ---- */
/* coefficientMatrix, CoefficientMatrix.f90:10-71 */
int biocfd_coefficient_matrix(int blk, int is_coarse);
/* non_uni_coeff, NavierStokes.f90:13-250 */
int biocfd_non_uni_coeff(int blk);

/* ---- momentum and boundary conditions ---- */
/* nsMomentum2order, NavierStokes.f90:252-912 */
int biocfd_ns_momentum(int blk, double dt, double re);
/* nsMomentum2order on blocks first..last (the reference calls it per block in
 * an OpenMP loop, main.F90:133-148; each block's step reads and writes only
 * its own arrays, so the calls are independent): the thin-surface pre-pass
 * per block, the dense pass of all blocks in one launch over a table of the
 * blocks (the default march shape, FP64 or the FP32 tendency; another
 * momentum variant runs block by block), then each block's thin-surface
 * pass.  The same bits as biocfd_ns_momentum on each block. */
int biocfd_ns_momentum_set(int first, int last, double dt, double re);
/* Dense momentum kernel of biocfd_ns_momentum (no reference counterpart):
 * variant 0 = the three components per thread, one workgroup per 64 x 4
 * cells of a plane; 1 = one launch per velocity component; 2 = as 0 with its
 * registers capped for four waves per SIMD; 3..8 = a workgroup marching a
 * 64 x TJ column tile up KC planes with the planes staged in LDS, (TJ, KC) =
 * (4, 16), (4, 32) (the default, 4), (8, 16), (8, 32), (4, 32) with
 * registers capped for three waves per SIMD, (4, 8).  pow2_products (default 1): on a block whose
 * every axis has one power-of-two spacing, the divisions by grid spacings
 * are products by their exact reciprocals (the same correctly rounded
 * values).  All settings give the same bits. */
int biocfd_momentum_config(int variant, int pow2_products);
/* Precision of the dense momentum pass (no reference counterpart; the
 * mixed-precision step of the precision recipe): 0 (default) = FP64, the
 * reference's arithmetic; 1 = the tendency dt*(-convection + diffusion/re -
 * grad p) in FP32 (FP32 copies of the planes, upwind tables and spacings,
 * the marching kernel's 64 x 4 x 32 shape) added to the FP64 velocity in
 * FP64.  The deferred cells of the thin-surface pass stay FP64.  Mode 1
 * changes the results at FP32 rounding level. */
int biocfd_momentum_precision(int fp32);
/* Divisions of the dense momentum march (no reference counterpart):
 * reciprocal_rows = 1 (default) divides by the upwind denominators ca6 hA,
 * 2 ck6 hB and by re through two FMA corrections of the product with their
 * reciprocals rounded once (tables made by biocfd_non_uni_coeff), which is
 * the correctly rounded quotient while the numerator lies in [2^-957, 2^700)
 * or is zero; a cell with another numerator is evaluated with divisions.
 * 0 = divisions everywhere.  Both give the same bits. */
int biocfd_momentum_divisions(int reciprocal_rows);
/* The deferred cells of nsMomentum2order's thin-surface pass
 * (NavierStokes.f90:310-490, 860-907) (no reference counterpart):
 * parallel = 1 (the default) evaluates them all at once, each reading the
 * ghost substitutions and restores the reference's in-order loop would have
 * left at that point (fixed by positions alone), then leaves the loop's final
 * state; 0 runs them in list order in one workgroup.  Same bits either way. */
int biocfd_ts_pass_config(int parallel);
/* velocityBC, BoundaryConditions.f90:11-74 */
int biocfd_velocity_bc(int blk, double dt, double uc);
/* solidCellBC, BoundaryConditions.f90:76-92 */
int biocfd_solid_cell_bc(int blk);
/* solidCellBC_move, BoundaryConditions.f90:94-115 */
int biocfd_solid_cell_bc_move(int blk, int64_t move_check);
/* updateVelocity_newv, PcorVcor.F90:379-397 */
int biocfd_update_velocity_newv(int blk, int64_t move_check);

/* ---- pressure correction ---- */
/* poissonSolver, PcorVcor.F90:22-220, over blocks first..last (1-based)
 * with the interfaces registered by biocfd_interface_create */
int biocfd_poisson_solver(int first, int last, int64_t pc_ita_max, double epsi, double dt,
                          double uc);
/* the pieces of poissonSolver, for unit tests: computeDiv :222-243,
 * REDBLACKSOR_linear :288-377, correctPressure+correctVelocity :245-286 */
int biocfd_compute_div(int blk);
int biocfd_redblack_sor(int blk, int64_t pc_ita_max, double dt, double epsi);
int biocfd_correct(int blk, double dt);
/* Kernel variant of the SOR iterations (no reference counterpart; the
 * reference has one loop nest).  variant 0: one launch per colour on
 * colour-split arrays; variant >= 1: one fused red+black launch per iteration
 * over tiles of tile_i x tile_j cells and `planes_per_workgroup` z-planes.
 * A fused variant may run two iterations per pass (temporal blocking); every
 * variant gives the same bits, the choice only changes speed. */
int biocfd_sor_config(int variant, int planes_per_workgroup);
/* The FP64 SOR kernel of one block's own SOR calls (no reference
 * counterpart): variant and planes per workgroup as biocfd_sor_config, for
 * the calls poisson_solver makes on that block alone (the coarse block; a
 * batched set of fine blocks keeps the library's variant); variant -1
 * returns the block to the library's choice.  Every variant gives the same
 * bits. */
int biocfd_sor_block_config(int blk, int variant, int planes_per_workgroup);
/* Stopping test of the FP64 SOR calls (rule 1 has no reference counterpart).
 * rule 0 (default): the reference's, max |pc - pco| < epsi over the fluid
 * cells tested every fifth iteration (PcorVcor.F90:351-375).  rule 1: the
 * FP64 residual max |b/dt - A pc| < epsi over the fluid cells, tested every
 * `cadence` (>= 1) iterations and after the last iteration of the budget;
 * derr2 then holds the residual of the last test.  nIterPcor accumulation
 * and the pcItaMax exit are the reference's under both rules.  Rule 1 needs
 * a fused SOR variant and runs the fine blocks' calls one after the other. */
int biocfd_sor_stop_rule(int rule, int cadence);
/* Batching of the fine blocks' SOR calls in poisson_solver (no reference
 * counterpart; the reference runs them one block per OpenMP thread,
 * PcorVcor.F90:86-129): on (the default), the calls of all fine blocks that
 * start at the same iteration advance together, one launch per pass over a
 * table of the blocks (pair-owned kernels, FP64); off, one block after the
 * other.  Both give the same bits. */
int biocfd_sor_batching(int on);
/* HIP graphs of the FP64 SOR passes of a single-block call (no reference
 * counterpart): on (the default), each run of passes between two host
 * polls whose pattern (check iterations, pass parity) was seen before in the
 * call's block replays a graph captured the first time, one graph launch
 * instead of the passes' and end kernels' launches; off, every pass is
 * launched.  Both give the same bits. */
int biocfd_sor_graphs(int on);
/* The coarse/fine interface updates inside poisson_solver (no reference
 * counterpart; the reference updates one interface after the other,
 * CoarseInterp.f90:447-502, FineInterp_bound.f90:554-603): on (the
 * default), the same update of every live interface and component runs in
 * one launch (per axis pass of the prolongation) from a task table, where
 * no interface reads what another writes (each fine block on one interface,
 * no block on both sides, the restrictions' coarse cells disjoint); else,
 * and off, interface by interface.  Both give the same bits. */
int biocfd_interp_batching(int on);
/* Graphs captured and graph launches of biocfd_sor_graphs since the
 * library was loaded (no reference counterpart; a test and bench probe). */
int biocfd_sor_graph_stats(int64_t *captures, int64_t *launches);
/* info = {tile_i, tile_j, threads per workgroup, z-plane prefetch depth,
 * workgroups per CU its register budget is compiled for, iterations per
 * pass, tile family (0: per-phase cell lists, 1: pair-owned cells)} of a
 * variant (zeros and 1 for variant 0); fails past the last one */
int biocfd_sor_variant_info(int variant, int info[7]);
/* Autotuning probe (no reference counterpart): time `iters` SOR iterations
 * of a variant on block `blk`'s current pc and b (milliseconds, HIP events);
 * pc and the block's SOR state are restored afterwards. */
int biocfd_sor_time(int blk, int variant, int planes_per_workgroup, int iters, double dt, double *ms);

/* ---- mixed precision (no reference counterpart: the reference solves in
 * FP64 only).  mode 1 replaces every REDBLACKSOR_linear call of
 * poissonSolver by FP64 defect correction with FP32 red-black SOR sweeps:
 * r = b/dt - A pc (FP64), `sweeps_per_correction` FP32 sweeps on A e = r from
 * e = 0, pc += e (FP64), until max|r| < residual_tol or the call's sweep
 * budget (pcItaMax + 1 - nIterPcor) is spent; nIterPcor counts sweeps and
 * derr2 holds max|r|.  mode 0 (default) is the FP64 method. */
int biocfd_set_precision(int mode);
/* sweeps_per_correction: 1..1024 */
int biocfd_mixed_config(int variant, int planes_per_workgroup, int sweeps_per_correction,
                        double residual_tol);
int biocfd_mixed_variant_info(int variant, int info[7]);
/* time `iters` FP32 sweeps of a variant on the residual of block `blk`'s pc */
int biocfd_mixed_time(int blk, int variant, int planes_per_workgroup, int iters, double dt, double *ms);
/* max over fluid cells of |b/dt - A pc| for block `blk` (FP64 diagnostic) */
int biocfd_residual_norm(int blk, double dt, double *out);

/* ---- adjoint (no reference counterpart: the reference has no adjoint) ----
 * FP64, geometry frozen (tags, index sets and interface tables as the
 * forward's).  Host arrays in the block layout ((nx+2)(ny+2)(nz+2) doubles).
 *
 * biocfd_projection: on block blk, b = pc = 0, nIterPcor = 0, computeDiv,
 * n_iter red-black SOR iterations of the FP64 method, correctPressure and
 * correctVelocity (PcorVcor.F90:67-77, 222-377) in place on ut, vt, wt, p;
 * no interface updates and no boundary conditions.
 * biocfd_projection_adjoint: its transpose with respect to (ut, vt, wt):
 * given the adjoints of (ut, vt, wt, p) after it, the adjoints of (ut, vt,
 * wt) before it (the adjoint of p passes unchanged): the transposed
 * correction, n_iter transposed red-black iterations in reverse order
 * (black then red, each with the neighbours' coefficients) and the
 * transposed divergence. */
int biocfd_projection(int blk, int64_t n_iter, double dt);
int biocfd_projection_adjoint(int blk, int64_t n_iter, double dt, const double *ubar, const double *vbar,
                              const double *wbar, const double *pbar, double *utbar, double *vtbar,
                              double *wtbar);
/* Transposes of the trilinear interface updates of one component (0 p / pc,
 * 1 u / ut, 2 v / vt, 3 w / wt) on interface itf (CoarseInterp.f90:447-502,
 * FineInterp_bound.f90:554-603).  coarse: on entry the adjoint of the
 * coarse field after coarseUpdate, on return before it (its targets
 * zeroed); fine: the restriction's transpose added.  fine_bd: on entry the
 * adjoint of the fine field after fineUpdate_*_bd, on return before it;
 * coarse: the prolongation's transpose added. */
int biocfd_coarse_update_adjoint(int itf, int comp, double *lam_coarse, double *lam_fine);
int biocfd_fine_update_bd_adjoint(int itf, int comp, double *lam_fine, double *lam_coarse);

/* The tangent-linear and adjoint time step of a single block with a body
 * (no reference counterpart), FP64, geometry frozen: IB/solid cells, nearest
 * elements, distances, normals and image-point cubes are the forward's at
 * the moment of the call, and the momentum stencil is differentiated at the
 * block's current u, v, w, p, including the thin-surface (TS) pass's ghost
 * substitutions and restores.  The sensitivities live on the device per
 * block: one array per field BF_U .. BF_P (u, v, w, ut, vt, wt, p) and per
 * TS ghost array (Blocks.f90:55-58), tangents for the tlm calls, adjoints
 * for the adj calls; the adj calls add the kinematic adjoint (the gradient with
 * respect to the kin vector of the forcing and stress calls) into a device
 * accumulator that biocfd_ad_kin reads.
 *
 * Forward pieces differentiated (a step of main.F90:133-263 in the drivers'
 * order): nsMomentum2order (NavierStokes.f90:252-912), velocityBC
 * (BoundaryConditions.f90:11-74), solidCellBC (:76-92), velocityForcing1,
 * pressureForcing1, velocityForcingField, pressureForcingField
 * (Forcing.f90:14-1075), poissonSolver on one block with a fixed SOR
 * count n_iter = pcItaMax + 1 (epsi 0; PcorVcor.F90:22-220), and the force
 * sums of stress_calculation (StressCalc.f90:14-208).  routine: 0
 * velocityForcing1, 1 pressureForcing1, 2 velocityForcingField, 3
 * pressureForcingField, 4 velocityForcingGhost, 5 pressureForcingGhost.
 * Step halves: start = momentum, velocityBC, solidCellBC, velocityForcing1,
 * velocityBC, poissonSolver, pressureForcing1 (at the geometry and
 * kinematics of the step's start); end = velocityForcingField,
 * pressureForcingField, velocityForcingGhost, pressureForcingGhost (after
 * the body moved). */
int biocfd_ad_reset(int blk);                      /* zero the sensitivities and the kin accumulator */
int biocfd_ad_upload(int blk, int field, const double *host);
int biocfd_ad_download(int blk, int field, double *host);
int biocfd_ad_kin(int blk, double kin_bar[BK_N], int clear);
/* the sensitivities of TS ghost array `which` (BGH_*), first `count` entries */
int biocfd_ad_ghost_upload(int blk, int which, const double *host, int64_t count);
int biocfd_ad_ghost_download(int blk, int which, double *host, int64_t count);
int biocfd_tlm_momentum(int blk, double dt, double re);
int biocfd_adj_momentum(int blk, double dt, double re);
int biocfd_tlm_velocity_bc(int blk, double dt, double uc);
int biocfd_adj_velocity_bc(int blk, double dt, double uc);
int biocfd_tlm_solid_cell_bc(int blk);
int biocfd_adj_solid_cell_bc(int blk);
int biocfd_tlm_forcing(int blk, int routine, const double kin[BK_N], const double kin_dot[BK_N], int rec);
int biocfd_adj_forcing(int blk, int routine, const double kin[BK_N], int rec);
int biocfd_tlm_poisson(int blk, int64_t n_iter, double dt, double uc);
int biocfd_adj_poisson(int blk, int64_t n_iter, double dt, double uc);
/* the four element sums stress_calculation divides by its projected areas:
 * viscous x, pressure x, viscous y, pressure y (StressCalc.f90:173-179) */
int biocfd_stress_sums(int blk, double mu_f, double rho_f, const double kin[BK_N], double out[4]);
int biocfd_tlm_stress(int blk, double mu_f, double rho_f, const double kin[BK_N], const double kin_dot[BK_N],
                      int geom, double out[4]);
int biocfd_adj_stress(int blk, double mu_f, double rho_f, const double kin[BK_N], const double sbar[4], int geom);
int biocfd_tlm_step_start(int blk, int64_t n_iter, double dt, double re, double uc, const double kin[BK_N],
                          const double kin_dot[BK_N], int rec);
int biocfd_tlm_step_end(int blk, const double kin[BK_N], const double kin_dot[BK_N], int rec);
int biocfd_adj_step_end(int blk, const double kin[BK_N], int rec);
int biocfd_adj_step_start(int blk, int64_t n_iter, double dt, double re, double uc, const double kin[BK_N],
                          int rec);
/* The moving geometry (the same calls with rec >= 0 / geom 1): the element
 * centroids, normals and lengths (computeSurfaceNorm, Search.F90:210-258)
 * enter every reconstruction (its normal distance, image point, cube
 * weights and surface term) and the force sums; the discrete structure (IB
 * cells, nearest elements, cubes, TS ghost choices) stays the forward's.
 * biocfd_ad_record(blk, slot): the forcing calls that follow keep what they
 * read (the interpolated arrays at the call's start, the in-place routines'
 * task outputs) in record `slot` (-1 stops); the tlm/adj calls with rec =
 * slot use them.  biocfd_ad_geometry_tangent sets the element tangents from
 * node-position tangents (n_nodes each, at the block's current node
 * positions; inor the normal orientation of computeSurfaceNorm); the adj
 * calls add element adjoints, which biocfd_ad_geometry_adjoint turns into
 * node-position adjoints (and clears with clear = 1). */
int biocfd_ad_record(int blk, int slot);
int biocfd_ad_geometry_tangent(int blk, double inor, const double *xdot, const double *ydot, const double *zdot);
int biocfd_ad_geometry_adjoint(int blk, double inor, double *xbar, double *ybar, double *zbar, int clear);
/* Step checkpoints: the device state a time step reads and writes (fields
 * u .. p, tags, IB / TS / solid lists, element and node data, counts) copied
 * into slot `slot` (0..4095, device memory) and back; free: slot -1 = all. */
int biocfd_snapshot_save(int blk, int slot);
int biocfd_snapshot_restore(int blk, int slot);
int biocfd_snapshot_free(int blk, int slot);

/* ---- immersed-boundary geometry, search and tagging (Search.F90) ---- */
/* computeSurfaceNorm :210-258 (nodes are the moved positions xnode1..) */
int biocfd_compute_surface_norm(int blk, const double *xnode1, const double *ynode1,
                                const double *znode1, int64_t inor);
/* tagging_th_core :260-386 */
int biocfd_tagging(int blk);
/* selectiveRetagging_th :520-651 */
int biocfd_selective_retagging(int blk, int64_t blk_mv_tag);
/* cellCount_solid :653-741 (counts + device lists) */
int biocfd_cell_count_solid(int blk);
/* cellCount_solid_coarse :1463-1532 */
int biocfd_cell_count_solid_coarse(int blk);
/* fine_block_cell :863-914, for the coarse block of every registered interface */
int biocfd_fine_block_cell(int coarse_blk);
/* computeNormDistance :743-861 */
int biocfd_compute_norm_distance(int blk);
/* findTScells :434-518 */
int biocfd_find_ts_cells(int blk);

/* ---- sharp-interface reconstruction (Forcing.f90) ---- */
int biocfd_velocity_forcing1(int blk, const double kin[BK_N]);      /* :92-360 */
int biocfd_pressure_forcing1(int blk, const double kin[BK_N]);      /* :14-90 */
int biocfd_velocity_forcing_field(int blk, const double kin[BK_N]); /* :813-1075 */
int biocfd_pressure_forcing_field(int blk, const double kin[BK_N]); /* :737-811 */
int biocfd_velocity_forcing_ghost(int blk, const double kin[BK_N]); /* :450-734 */
int biocfd_pressure_forcing_ghost(int blk, const double kin[BK_N]); /* :362-448 */
/* stress_calculation, StressCalc.f90:14-208: out = {viscous_drag,
 * pressure_drag, viscous_lift, pressure_lift} coefficients */
int biocfd_stress_calculation(int blk, double mu_f, double rho_f, const double kin[BK_N],
                              double out[4]);
/* biocfd_stress_calculation without any device allocation (no counterpart in
 * the reference): for callers that cannot allocate, such as code under HIP
 * graph capture or a tight serving loop.  Its scratch (per-element terms,
 * the coefficients) is a static device array of the library holding up to
 * 65536 elements (3 MB per device, kept until the process ends); a body with
 * more elements is refused with an error.  Same arguments, kernels and
 * results as biocfd_stress_calculation.  Every call on one device shares
 * that scratch: calls are serialised by the library, and calls that
 * overlapped on one device would race. */
int biocfd_stress_calculation_noalloc(int blk, double mu_f, double rho_f, const double kin[BK_N],
                                      double out[4]);

/* ---- coarse/fine interfaces ---- */
/* register interface `itf` (1-based) between coarse block a and fine block b
 * and build its index tables (interfaceDetail / set_interface_detail,
 * InterfaceDetail.f90:10-241) */
int biocfd_interface_create(int itf, int a_blk, int b_blk, int64_t a_msh, int64_t b_msh,
                            const double extent[6]);
/* which: 0 = p,u,v,w ; 1 = ut,vt,wt ; 2 = pc (+pco) */
/* coarseUpdate / coarseUpdate_newv / coarseUpdate_pc, CoarseInterp.f90:16-437 */
int biocfd_coarse_update(int which);
/* fineUpdate_bd_mv (which 0) / fineUpdate_newv_bd (1) / fineUpdate_pc_bd (2),
 * FineInterp_bound.f90:17-544 */
int biocfd_fine_update_bd(int which);
/* the same on interface itf alone (fineUpdate_bd_mv(g) with which 0,
 * FineInterp_bound.f90:311-544) */
int biocfd_fine_update_bd_one(int itf, int which);
int biocfd_interface_table(int itf, int comp, int axis, int64_t *host, int64_t *counter, int64_t cap);

/* ---- body kinematics (host arithmetic on host arrays, Search.F90:21-208) ---- */
/* state: {xshift,yshift,zshift,a0,yamp,bfreq, xpth1,xpth2,ypth2,xchg,ychg,yt,ydot,yddot,xt,xdot,
 *         thetaDot,thetaDDot,thetaDot1,thetaDDot1,thetaDot2,thetaDDot2,alphaDot,
 *         piv_x,piv_y,piv_z,xmove,ymove,zmove,inity_cent,nxty_cent,initx_cent,nxtx_cent,
 *         initz_cent,nxtz_cent} -- 35 doubles, updated in place */
enum { BKS_N = 35 };
/* shiftSurfaceNodesInitial :49-124 */
int biocfd_shift_surface_nodes_initial(int64_t n_nodes, const double *xnode, const double *ynode,
                                       const double *znode, const int64_t *node_id, double aoa,
                                       double piv_pt, int64_t ita, double dt, double dxmin,
                                       double totime, double *state, double *xnode1,
                                       double *ynode1, double *znode1);
/* computeSurfaceVariables :126-208 (without its per-step log file) */
int biocfd_compute_surface_variables(int64_t n_nodes, const double *xnode, const double *ynode,
                                     const double *znode, const int64_t *node_id,
                                     double phase_angle, double piv_pt, int64_t ita, double dt,
                                     double dxmin, double freq, double totime, double *state,
                                     double *xnode1, double *ynode1, double *znode1);

/* ---- block move (Search.F90:21-47, :415-432, :916-1461; FineInterp.f90:10-281) ---- */
/* findDistnode :21-47: the 1-based marker nodes {mk, mkx1, mkx2} among the
 * ID-51 nodes (body-frame coordinates), 0 where no ID-51 node qualifies;
 * block_move_check needs all three */
int biocfd_find_dist_node(int64_t n_nodes, const double *xnode, const double *znode,
                          const int64_t *node_id, int64_t markers[3]);
/* block_move_check :997-1203 for interface itf: when the body of its fine
 * block (nodes xnode1/ynode1 at the markers, kinematic state) is within the
 * margins of the interface, the block moves: *moved = 1 (the caller sets the
 * block's move_check and blk_mv_tag to 1), *coarse_flcnt_check = 1, the
 * state's init*_cent are updated, the interface extents shift and the fine
 * block's u, v, w, p are kept for the shifted copy; else *moved = 0 */
int biocfd_block_move_check(int itf, const double *xnode1, const double *ynode1,
                            const int64_t markers[3], double *state, int64_t *moved,
                            int64_t *coarse_flcnt_check);
/* change_block_coords :1208-1264: the block's grids shifted by its last move
 * (nothing unless move_check == 1) */
int biocfd_change_block_coords(int blk, int64_t move_check);
/* change_block_coords_interfaces :1269-1370: interior mask cell_n of the
 * moved fine block, its kept overlap copied to the new indices, zero
 * elsewhere (move_check of the interface's fine block) */
int biocfd_change_block_coords_interfaces(int itf, int64_t move_check);
/* change_block_interface :1372-1461: coarse indices of the tables shifted */
int biocfd_change_block_interface(int itf, int64_t move_check);
/* cellCount_solid_coarse_mv :916-995: the coarse block's lists rebuilt when
 * *coarse_flcnt_check == 1; the flag is cleared */
int biocfd_cell_count_solid_coarse_mv(int coarse_blk, int64_t *coarse_flcnt_check);
/* fineUpdate_mv, FineInterp.f90:10-281: coarse -> fine volume fill of the
 * cells of interface itf's fine block outside its interior mask */
int biocfd_fine_update_mv(int itf);
/* tagging_th_move :415-432: nothing unless move_check == 1; else the search
 * box's tagging, then fineUpdate_mv and fineUpdate_bd_mv of every interface
 * whose fine block is blk */
int biocfd_tagging_move(int blk, int64_t move_check);
/* a block's grid array (enum biocfd_grid) back to the host, n entries */
int biocfd_grid_download(int blk, int which, double *host, int64_t n);

/* ---- output sync (WriteOutput_corner1.F90:80-139) ---- */
/* the state the field output writes, per interior cell (nx*ny*nz entries,
 * i fastest): cell-centre velocities u1 = 0.5 (u(i) + u(i-1)), v1, w1, the
 * pressure and the tags cell, cell_n, cell_pr (write_output_ascii :97-107),
 * assembled on the device and copied to the host in one transfer */
int biocfd_output_fields(int blk, double *u1, double *v1, double *w1, double *p, int64_t *cell,
                         int64_t *cell_n, int64_t *cell_pr);

#ifdef __cplusplus
}
#endif
#endif
