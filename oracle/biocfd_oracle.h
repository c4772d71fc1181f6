/*
 * Test infrastructure, not product code: a plain-C (FP64, serial) restatement
 * of the reference Bio-CFD time step, used only by tests/, the smoke check and
 * bench.py's CPU-baseline leg as the checker.
 *
 * Arrays keep the reference's Fortran layout: column-major, 1-based indices
 * with one ghost layer, i.e. field(i,j,k) for i=1..nx+2 lives at
 * f[(i-1) + (nx+2)*((j-1) + (ny+2)*(k-1))].  Tag arrays cell2/cell_pr/nodeIdTag
 * have one extra entry per axis ((nx+3)^3), as allocated in the reference
 * (src/AllocateArrays.f90:40-45).  Index lists are (count,3) column-major.
 *
 * Parity status of each routine (what pins it) is listed in
 * oracle/README.md.
 */
#ifndef BIOCFD_ORACLE_H
#define BIOCFD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_NGHOST = 14 };
/* order of the TS ghost arrays (Blocks.f90:55-58) */
enum { G_P = 0, G_PT, G_U1, G_U1T, G_U2, G_U2T, G_V1, G_V1T, G_V2, G_V2T, G_W1, G_W1T, G_W2, G_W2T };
/* order of the 3rd-order upwind tables (Blocks.f90:29-31) */
enum { T_UU = 0, T_VV, T_WW, T_UV, T_UW, T_VU, T_VW, T_WU, T_WV, T_COUNT };

typedef struct orc_block {
    int64_t nx, ny, nz, cintp, has_ib, moving;
    double dx, dy, dz, omega;
    double *x1, *y1, *z1, *deltax, *deltay, *deltaz;
    double *xu, *yu, *zu, *xv, *yv, *zv, *xw, *yw, *zw, *xp, *yp, *zp;
    double *u, *v, *w, *ut, *vt, *wt, *p, *b, *pc, *pco;
    int64_t *cell, *cell_n;                 /* (n+2)^3 */
    int64_t *cell2, *cell_pr, *nodeIdTag;   /* (n+3)^3 */
    double *Acx, *Acy, *Acz;                /* (n,3) */
    double *ca[T_COUNT], *ck[T_COUNT];      /* (6, n+2) */
    int64_t fluidCount, solidCount, ibCount, redCount, blackCount, tsCount;
    int64_t *fluid, *red, *black, *solid, *intercepted, *ts, *index_ts;
    int64_t *nelp, *nelu1, *nelu2, *nelv1, *nelv2, *nelw1, *nelw2;
    double *pNormDis, *u1NormDis, *u2NormDis, *v1NormDis, *v2NormDis, *w1NormDis, *w2NormDis;
    double *ghost[ORC_NGHOST];
    int64_t ibNodes, ibElems;
    int64_t *ibNodeId, *elP1, *elP2, *elP3, *surfId;
    double *xnode, *ynode, *znode, *xnode1, *ynode1, *znode1;
    double *xcent, *ycent, *zcent, *cosA, *cosB, *cosG, *elen;
    int64_t search[6];
    double xshift, yshift, zshift, a0, yamp, bfreq;
    double xdot, ydot, yddot, xt, yt;
    double thetaDot, thetaDDot, thetaDot1, thetaDDot1, thetaDot2, thetaDDot2, alphaDot;
    double piv_x, piv_y, piv_z;
    double xpth1, xpth2, ypth2, xchg, ychg, xmove, ymove, zmove;
    double inity_cent, nxty_cent, initx_cent, nxtx_cent, initz_cent, nxtz_cent;
    int64_t nIterPcor, mk, mkx1, mkx2, move_check, blk_mv_tag;
    /* block move (Blocks.f90:64-71): shift in cells, the copied index ranges
     * before (cpy[0..5]) and after (cpy_mv[0..5]) the move, x/y/z start/end */
    int64_t move_amtx, move_amty, move_amtz, cpy[6], cpy_mv[6];
    double *u_dum, *v_dum, *w_dum, *p_dum;
    double derr1, derr2, derrStdSt;
    double viscous_drag, pressure_drag, viscous_lift, pressure_lift;
} orc_block;

typedef struct orc_interface {
    int64_t a_blk, b_blk, a_msh, b_msh;     /* 1-based block numbers */
    double xs, xe, ys, ye, zs, ze;
    double xs_new, xe_new, ys_new, ye_new, zs_new, ze_new;   /* overlap kept by a move */
    /* det tables (3, cap) column-major and counters, order p,u,v,w x,y,z */
    int64_t *det[4][3];
    int64_t counter[4][3];
    int64_t cap[3];
} orc_interface;

typedef struct orc_case {
    int64_t nblocks, nintf, ib_block1, pcItaMax, nsteps, inor, seed;
    double dt, re, uc, epsi, noise_amp, aoa, phase_angle_deg, piv_pt, dxmin, freq, mu_f, rho_f;
    orc_block *blk;          /* max(nblocks,2) entries; index 0 = block 1 */
    orc_interface *intf;
    int64_t ndump, *dump_steps;
} orc_case;

/* -------- case driver (mirrors the reference driver's call order) -------- */
int orc_case_load(orc_case *c, const char *path);
void orc_case_free(orc_case *c);
int orc_case_setup(orc_case *c);
int orc_case_step(orc_case *c, int64_t ita, double *totime, double *totalTime);
int orc_case_dump(const orc_case *c, const char *out_dir, int64_t step);
/* run a case end to end; when seconds != NULL the time loop is timed and no
 * dumps are written except the final step's */
int orc_run_case(const char *case_path, const char *out_dir, int64_t nsteps_override,
                 double *seconds);

/* -------- per-routine restatements (reference procedure in brackets) -------- */
void orc_coefficient_matrix(orc_block *b, int is_coarse);      /* coefficientMatrix */
void orc_non_uni_coeff(orc_block *b);                          /* non_uni_coeff */
void orc_ns_momentum(orc_block *b, double dt, double re);      /* nsMomentum2order */
void orc_velocity_bc(orc_block *b, double dt, double uc);      /* velocityBC */
void orc_solid_cell_bc(orc_block *b);                          /* solidCellBC */
void orc_solid_cell_bc_move(orc_block *b);                     /* solidCellBC_move */
void orc_update_velocity_newv(orc_block *b);                   /* updateVelocity_newv */
void orc_compute_div(orc_block *b);                            /* computeDiv */
void orc_redblack_sor(orc_block *b, int64_t pcItaMax, double dt, double epsi);
void orc_correct_pressure(orc_block *b);
void orc_correct_velocity(orc_block *b, double dt);
void orc_poisson_solver(orc_case *c, int64_t ita, double totime, double *totalTime);
void orc_find_dist_node(orc_block *b);
void orc_shift_surface_nodes_initial(orc_block *b, double aoa, double piv_pt, int64_t ita,
                                     double dt, double dxmin, double totime);
void orc_compute_surface_variables(orc_block *b, double phase_angle, double piv_pt,
                                   int64_t ita, double dt, double dxmin, double freq,
                                   double totime);
void orc_compute_surface_norm(orc_block *b, int64_t inor);
void orc_tagging_core(orc_block *b);
void orc_selective_retagging(orc_block *b);
void orc_cell_count_solid(orc_block *b);
void orc_cell_count_solid_coarse(orc_block *b);
void orc_fine_block_cell(orc_case *c);
void orc_compute_norm_distance(orc_block *b);
void orc_find_ts_cells(orc_block *b);
void orc_pressure_forcing1(orc_block *b);
void orc_velocity_forcing1(orc_block *b);
void orc_pressure_forcing_field(orc_block *b);
void orc_velocity_forcing_field(orc_block *b);
void orc_pressure_forcing_ghost(orc_block *b);
void orc_velocity_forcing_ghost(orc_block *b);
void orc_stress_calculation(orc_block *b, double mu_f, double rho_f);
void orc_interface_detail(orc_case *c);
void orc_coarse_update(orc_case *c);
void orc_coarse_update_newv(orc_case *c);
void orc_coarse_update_pc(orc_case *c);
void orc_fine_update_pc_bd(orc_case *c);
void orc_fine_update_newv_bd(orc_case *c);
void orc_fine_update_bd_mv(orc_case *c, int64_t g);
void orc_fine_update_mv(orc_case *c, int64_t g);               /* fineUpdate_mv */

/* -------- block move (Search.F90:916-1461) -------- */
int orc_move_markers_defined(const orc_case *c);
void orc_block_move_check(orc_case *c, int64_t *coarse_flcnt_check);
void orc_change_block_coords(orc_block *b);
void orc_change_block_coords_interfaces(orc_interface *it, orc_block *b);
void orc_change_block_interface(orc_interface *it, const orc_block *b);
void orc_cell_count_solid_coarse_mv(orc_case *c, int64_t *coarse_flcnt_check);
void orc_tagging_move(orc_case *c, int64_t g);                 /* tagging_th_move */

/* scalar helpers (Interpolation.f90, Forcing.f90:1080-1183) */
double orc_linear_interpolation(double x, double x0, double x1, double y0, double y1);
double orc_bilinear_interpolation(double x, double y, double x1, double x2, double y1,
                                  double y2, const double fvals[4]);
int64_t orc_find_index_in_array(double value, const double *array1, int64_t start,
                                int64_t finish, int *found);
double orc_noise_u(int64_t seed, int64_t field_id, int64_t lin);

#ifdef __cplusplus
}
#endif
#endif
