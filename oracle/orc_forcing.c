/*
 * Test infrastructure: FP64 restatement of the sharp-interface IB
 * reconstruction, reference src/Forcing.f90 (pressureForcing1 :14-90,
 * velocityForcing1 :92-360, pressureForcingGhost :362-448,
 * velocityForcingGhost :450-734, pressureForcingField :737-811,
 * velocityForcingField :813-1075, compute_derivative :1080-1089,
 * compute_value_and_derivatives :1091-1161, find_index_in_array :1165-1183),
 * src/Interpolation.f90:16-42 and src/StressCalc.f90:14-208.
 * Cells are processed in list order with in-place writes, as the serial CPU
 * build does.
 */
#include <stdio.h>
#include "orc_internal.h"

static long g_find_miss = 0;

double orc_linear_interpolation(double x, double x0, double x1, double y0, double y1)
{
    return y0 + (x - x0) * (y1 - y0) / (x1 - x0);
}

double orc_bilinear_interpolation(double x, double y, double x1, double x2, double y1,
                                  double y2, const double f[4])
{
    double numer = f[0] * (x2 - x) * (y2 - y) + f[2] * (x2 - x) * (y - y1) +
                   f[1] * (x - x1) * (y2 - y) + f[3] * (x - x1) * (y - y1);
    double denom = (x2 - x1) * (y2 - y1);
    return numer / denom;
}

int64_t orc_find_index_in_array(double value, const double *a, int64_t start, int64_t finish,
                                int *found)
{
    for (int64_t i = start; i <= finish; ++i) {
        if (i < 1) continue; /* the reference would read before the array here */
        if (value >= V1(a, i) && value < V1(a, i + 1)) {
            if (found) *found = 1;
            return i;
        }
    }
    if (found) *found = 0;
    g_find_miss++;
    return start < 1 ? 1 : start;
}

static double cderiv(double x, double x2, double x1, double p_x, double p_x2, double p_x1)
{
    double h2 = fabs(x2 - x);
    double h1 = fabs(x1 - x);
    return (h1 * h1 * p_x2 - h2 * h2 * p_x1 + (h2 * h2 - h1 * h1) * p_x) /
           (h1 * h2 * (h1 + h2) + 1e-16);
}

/* compute_value_and_derivatives (Forcing.f90:1091-1161) */
static void value_and_derivs(const orc_block *b, double x, double y, double z, int64_t i,
                             int64_t j, int64_t k, const double *xg, const double *yg,
                             const double *zg, int off, const double *var, double *val,
                             double d[3])
{
    int64_t ii = i, jj = j, kk = k;
    if (off == 1) ii = i - 1;
    if (off == 2) jj = j - 1;
    if (off == 3) kk = k - 1;
#define VR(a, c, e) var[IX(b, a, c, e)]
    double f11[4] = { VR(ii, jj, kk), VR(ii, jj + 1, kk), VR(ii, jj, kk + 1), VR(ii, jj + 1, kk + 1) };
    double f12[4] = { VR(ii + 1, jj, kk), VR(ii + 1, jj + 1, kk), VR(ii + 1, jj, kk + 1), VR(ii + 1, jj + 1, kk + 1) };
    double f21[4] = { VR(ii, jj, kk), VR(ii + 1, jj, kk), VR(ii, jj, kk + 1), VR(ii + 1, jj, kk + 1) };
    double f22[4] = { VR(ii, jj + 1, kk), VR(ii + 1, jj + 1, kk), VR(ii, jj + 1, kk + 1), VR(ii + 1, jj + 1, kk + 1) };
    double f31[4] = { VR(ii, jj, kk), VR(ii + 1, jj, kk), VR(ii, jj + 1, kk), VR(ii + 1, jj + 1, kk) };
    double f32[4] = { VR(ii, jj, kk + 1), VR(ii + 1, jj, kk + 1), VR(ii, jj + 1, kk + 1), VR(ii + 1, jj + 1, kk + 1) };
#undef VR
    double t11 = orc_bilinear_interpolation(y, z, V1(yg, j), V1(yg, j + 1), V1(zg, k), V1(zg, k + 1), f11);
    double t12 = orc_bilinear_interpolation(y, z, V1(yg, j), V1(yg, j + 1), V1(zg, k), V1(zg, k + 1), f12);
    double t21 = orc_bilinear_interpolation(x, z, V1(xg, i), V1(xg, i + 1), V1(zg, k), V1(zg, k + 1), f21);
    double t22 = orc_bilinear_interpolation(x, z, V1(xg, i), V1(xg, i + 1), V1(zg, k), V1(zg, k + 1), f22);
    double t31 = orc_bilinear_interpolation(x, y, V1(xg, i), V1(xg, i + 1), V1(yg, j), V1(yg, j + 1), f31);
    double t32 = orc_bilinear_interpolation(x, y, V1(xg, i), V1(xg, i + 1), V1(yg, j), V1(yg, j + 1), f32);
    *val = orc_linear_interpolation(x, V1(xg, i), V1(xg, i + 1), t11, t12);
    d[0] = cderiv(x, V1(xg, i + 1), V1(xg, i), *val, t12, t11);
    d[1] = cderiv(y, V1(yg, j + 1), V1(yg, j), *val, t22, t21);
    d[2] = cderiv(z, V1(zg, k + 1), V1(zg, k), *val, t32, t31);
}

/* search limits for one image-point lookup: start/finish per axis and clamps */
typedef struct {
    int64_t s[3], f[3];
    int64_t clamp_hi_from[3], clamp_hi_to[3];   /* if idx == from -> to (0 = none) */
    int64_t clamp_lo_from[3], clamp_lo_to[3];
} lookup_t;

enum { K_P = 0, K_U, K_V, K_W };

/* one quadratic reconstruction; sgn = +1 (fluid side, *1 and *Field) or -1 (*Ghost) */
static double recon_velocity(orc_block *b, int comp, int64_t e, double sur2, double bx,
                             double by, double bz, int64_t i, int64_t j, int64_t k,
                             const lookup_t *lk, double sgn, double vsurf)
{
    double pt1 = 1.5 * sqrt(sq(V1(b->deltax, i)) + sq(V1(b->deltay, j)) + sq(V1(b->deltaz, k))) +
                 (fabs(sur2) - sur2) * 0.5;
    double px, py, pz;
    if (sgn > 0) {
        px = bx + pt1 * V1(b->cosA, e); py = by + pt1 * V1(b->cosB, e); pz = bz + pt1 * V1(b->cosG, e);
    } else {
        px = bx - pt1 * V1(b->cosA, e); py = by - pt1 * V1(b->cosB, e); pz = bz - pt1 * V1(b->cosG, e);
    }
    const double *xg, *yg, *zg, *var;
    switch (comp) {
    case K_U: xg = b->xu; yg = b->yu; zg = b->zu; var = b->ut; break;
    case K_V: xg = b->xv; yg = b->yv; zg = b->zv; var = b->vt; break;
    default: xg = b->xw; yg = b->yw; zg = b->zw; var = b->wt; break;
    }
    int64_t ix = orc_find_index_in_array(px, xg, lk->s[0], lk->f[0], NULL);
    int64_t iy = orc_find_index_in_array(py, yg, lk->s[1], lk->f[1], NULL);
    int64_t iz = orc_find_index_in_array(pz, zg, lk->s[2], lk->f[2], NULL);
    int64_t idx[3] = { ix, iy, iz };
    for (int a = 0; a < 3; ++a) {
        if (lk->clamp_hi_from[a] && idx[a] == lk->clamp_hi_from[a]) idx[a] = lk->clamp_hi_to[a];
        if (lk->clamp_lo_from[a] && idx[a] == lk->clamp_lo_from[a]) idx[a] = lk->clamp_lo_to[a];
    }
    double val, d[3];
    value_and_derivs(b, px, py, pz, idx[0], idx[1], idx[2], xg, yg, zg, comp, var, &val, d);
    double dn = d[0] * V1(b->cosA, e) + d[1] * V1(b->cosB, e) + d[2] * V1(b->cosG, e);
    if (sgn < 0) dn = -1.0 * dn;
    double n1 = pt1 + sur2;
    double cval = vsurf;
    double bval = 2.0 / n1 * (val - vsurf) - dn;
    double aval = dn / n1 - (val - vsurf) / (n1 * n1);
    return aval * (sur2 * sur2) + bval * sur2 + cval;
}

static double recon_pressure(orc_block *b, int64_t e, double sur2, int64_t i, int64_t j,
                             int64_t k, const lookup_t *lk, double sgn, double dpdn)
{
    double pt1 = 1.5 * sqrt(sq(V1(b->deltax, i)) + sq(V1(b->deltay, j)) + sq(V1(b->deltaz, k))) +
                 (fabs(sur2) - sur2) * 0.5;
    double px, py, pz;
    if (sgn > 0) {
        px = V1(b->xp, i) + pt1 * V1(b->cosA, e); py = V1(b->yp, j) + pt1 * V1(b->cosB, e);
        pz = V1(b->zp, k) + pt1 * V1(b->cosG, e);
    } else {
        px = V1(b->xp, i) - pt1 * V1(b->cosA, e); py = V1(b->yp, j) - pt1 * V1(b->cosB, e);
        pz = V1(b->zp, k) - pt1 * V1(b->cosG, e);
    }
    int64_t ix = orc_find_index_in_array(px, b->xp, lk->s[0], lk->f[0], NULL);
    int64_t iy = orc_find_index_in_array(py, b->yp, lk->s[1], lk->f[1], NULL);
    int64_t iz = orc_find_index_in_array(pz, b->zp, lk->s[2], lk->f[2], NULL);
    double val, d[3];
    value_and_derivs(b, px, py, pz, ix, iy, iz, b->xp, b->yp, b->zp, 0, b->p, &val, d);
    double dn = d[0] * V1(b->cosA, e) + d[1] * V1(b->cosB, e) + d[2] * V1(b->cosG, e);
    if (sgn < 0) dn = -1.0 * dn;
    double n1 = pt1 + sur2;
    double bval = dpdn;
    double aval = (dn - dpdn) / (2.0 * n1);
    double cval = val - (dn + dpdn) * n1 * 0.5;
    return aval * (sur2 * sur2) + bval * sur2 + cval;
}

/* acceleration terms of dp/dn for element e (Forcing.f90:29-52) */
static double dpdn_of(orc_block *b, int64_t e, int ghost)
{
    double ac_z = 0.0, ac_y = 0.0, at_z = 0.0, at_y = 0.0;
    int64_t sid = V1(b->surfId, e);
    if (sid == 50) {
        ac_z = 0.0; ac_y = 0.0; at_z = 0.0; at_y = 0.0;
    } else if (sid == 51 || sid == 52) {
        b->thetaDot = (sid == 51) ? b->thetaDot1 : b->thetaDot2;
        b->thetaDDot = (sid == 51) ? b->thetaDDot1 : b->thetaDDot2;
        ac_z = -sq(b->thetaDot) * (V1(b->zcent, e) - b->piv_z);
        ac_y = -sq(b->thetaDot) * (V1(b->ycent, e) - b->piv_y);
        at_z = b->thetaDDot * (V1(b->ycent, e) - b->piv_y);
        at_y = -b->thetaDDot * (V1(b->zcent, e) - b->piv_z);
    }
    if (!ghost)
        return -((ac_z + at_z) * V1(b->cosG, e) + (ac_y + at_y) * V1(b->cosB, e)) - b->yddot * V1(b->cosB, e);
    return ((ac_z + at_z) * V1(b->cosG, e) + (ac_y + at_y) * V1(b->cosB, e)) + (b->yddot * V1(b->cosB, e));
}

/* surface velocities (Forcing.f90:112-118, 194-204, 278-286 and the *Field /
 * *Ghost variants that add the alphaDot term to v) */
static double surf_u(orc_block *b, int64_t e) { (void)e; return b->xdot; }
static double surf_v(orc_block *b, int64_t e, int with_alpha)
{
    int64_t sid = V1(b->surfId, e);
    if (sid == 50) return b->ydot;
    b->thetaDot = (sid == 51) ? b->thetaDot1 : b->thetaDot2;
    if (!with_alpha) return -b->thetaDot * (V1(b->zcent, e) - b->piv_z) + b->ydot;
    return -b->thetaDot * (V1(b->zcent, e) - b->piv_z) + (-b->alphaDot) * (V1(b->xcent, e) - b->piv_x) + b->ydot;
}
static double surf_w(orc_block *b, int64_t e)
{
    int64_t sid = V1(b->surfId, e);
    if (sid == 50) return 0.0;
    b->thetaDot = (sid == 51) ? b->thetaDot1 : b->thetaDot2;
    return b->thetaDot * (V1(b->ycent, e) - b->piv_y);
}

static void lk_full(lookup_t *lk, int64_t s0, int64_t f0, int64_t s1, int64_t f1, int64_t s2, int64_t f2)
{
    for (int a = 0; a < 3; ++a) { lk->clamp_hi_from[a] = lk->clamp_lo_from[a] = 0; lk->clamp_hi_to[a] = lk->clamp_lo_to[a] = 0; }
    lk->s[0] = s0; lk->f[0] = f0; lk->s[1] = s1; lk->f[1] = f1; lk->s[2] = s2; lk->f[2] = f2;
}

static void lk_window(lookup_t *lk, int64_t i, int64_t j, int64_t k)
{
    lk_full(lk, i - 7, i + 7, j - 7, j + 7, k - 7, k + 7);
}

void orc_pressure_forcing1(orc_block *b)
{
    int64_t I = b->ibCount;
    for (int64_t n = 1; n <= I; ++n) {
        int64_t e = V1(b->nelp, n);
        double dpdn = dpdn_of(b, e, 0);
        int64_t i = LIST(b->intercepted, I, n, 1), j = LIST(b->intercepted, I, n, 2), k = LIST(b->intercepted, I, n, 3);
        lookup_t lk;
        lk_full(&lk, 1, b->nx + 1, 1, b->ny + 1, 1, b->nz + 1);
        FLD(b, p, i, j, k) = recon_pressure(b, e, V1(b->pNormDis, n), i, j, k, &lk, 1.0, dpdn);
    }
}

void orc_pressure_forcing_field(orc_block *b)
{
    int64_t I = b->ibCount;
    for (int64_t n = 1; n <= I; ++n) {
        int64_t e = V1(b->nelp, n);
        double dpdn = dpdn_of(b, e, 0);
        int64_t i = LIST(b->intercepted, I, n, 1), j = LIST(b->intercepted, I, n, 2), k = LIST(b->intercepted, I, n, 3);
        lookup_t lk;
        lk_window(&lk, i, j, k);
        FLD(b, p, i, j, k) = recon_pressure(b, e, V1(b->pNormDis, n), i, j, k, &lk, 1.0, dpdn);
    }
}

void orc_pressure_forcing_ghost(orc_block *b)
{
    int64_t T = b->tsCount;
    for (int64_t n = 1; n <= T; ++n) {
        int64_t i = LIST(b->ts, T, n, 1), j = LIST(b->ts, T, n, 2), k = LIST(b->ts, T, n, 3);
        int64_t t = V1(b->index_ts, n);
        int64_t e = V1(b->nelp, t);
        double dpdn = dpdn_of(b, e, 1);
        lookup_t lk;
        lk_window(&lk, i, j, k);
        V1(b->ghost[G_P], n) = recon_pressure(b, e, -V1(b->pNormDis, t), i, j, k, &lk, -1.0, dpdn);
        V1(b->ghost[G_PT], n) = FLD(b, p, i, j, k);
    }
}

void orc_velocity_forcing1(orc_block *b)
{
    int64_t I = b->ibCount, nx = b->nx, ny = b->ny, nz = b->nz;
    for (int64_t n = 1; n <= I; ++n) {
        int64_t i = LIST(b->intercepted, I, n, 1), j = LIST(b->intercepted, I, n, 2), k = LIST(b->intercepted, I, n, 3);
        lookup_t lk;
        int64_t e;
        /* u(i,j,k) */
        e = V1(b->nelu2, n);
        lk_full(&lk, 1, nx + 2, 1, ny + 1, 1, nz + 1);
        lk.clamp_hi_from[0] = nx + 2; lk.clamp_hi_to[0] = nx + 1;
        FLD(b, ut, i, j, k) = recon_velocity(b, K_U, e, V1(b->u2NormDis, n), V1(b->xu, i + 1), V1(b->yu, j),
                                             V1(b->zu, k), i, j, k, &lk, 1.0, surf_u(b, e));
        /* u(i-1,j,k) */
        e = V1(b->nelu1, n);
        lk_full(&lk, 1, nx + 1, 1, ny + 1, 1, nz + 1);
        lk.clamp_lo_from[0] = 1; lk.clamp_lo_to[0] = 2;
        FLD(b, ut, i - 1, j, k) = recon_velocity(b, K_U, e, V1(b->u1NormDis, n), V1(b->xu, i), V1(b->yu, j),
                                                 V1(b->zu, k), i, j, k, &lk, 1.0, surf_u(b, e));
        /* v(i,j,k) */
        e = V1(b->nelv2, n);
        double vs = surf_v(b, e, 0);
        lk_full(&lk, 1, nx + 1, 1, ny + 2, 1, nz + 1);
        lk.clamp_hi_from[1] = ny + 2; lk.clamp_hi_to[1] = ny + 1;
        FLD(b, vt, i, j, k) = recon_velocity(b, K_V, e, V1(b->v2NormDis, n), V1(b->xv, i), V1(b->yv, j + 1),
                                             V1(b->zv, k), i, j, k, &lk, 1.0, vs);
        /* v(i,j-1,k) */
        e = V1(b->nelv1, n);
        vs = (V1(b->surfId, e) == 50) ? 0.0 + b->ydot : surf_v(b, e, 0);
        lk_full(&lk, 1, nx + 1, 1, ny + 1, 1, nz + 1);
        lk.clamp_lo_from[1] = 1; lk.clamp_lo_to[1] = 2;
        FLD(b, vt, i, j - 1, k) = recon_velocity(b, K_V, e, V1(b->v1NormDis, n), V1(b->xv, i), V1(b->yv, j),
                                                 V1(b->zv, k), i, j, k, &lk, 1.0, vs);
        /* w(i,j,k) */
        e = V1(b->nelw2, n);
        double ws = surf_w(b, e);
        lk_full(&lk, 1, nx + 1, 1, ny + 1, 1, nz + 2);
        lk.clamp_hi_from[2] = nz + 2; lk.clamp_hi_to[2] = nz + 1;
        FLD(b, wt, i, j, k) = recon_velocity(b, K_W, e, V1(b->w2NormDis, n), V1(b->xw, i), V1(b->yw, j),
                                             V1(b->zw, k + 1), i, j, k, &lk, 1.0, ws);
        /* w(i,j,k-1) */
        e = V1(b->nelw1, n);
        ws = surf_w(b, e);
        lk_full(&lk, 1, nx + 1, 1, ny + 1, 1, nz + 1);
        lk.clamp_lo_from[2] = 1; lk.clamp_lo_to[2] = 2;
        FLD(b, wt, i, j, k - 1) = recon_velocity(b, K_W, e, V1(b->w1NormDis, n), V1(b->xw, i), V1(b->yw, j),
                                                 V1(b->zw, k), i, j, k, &lk, 1.0, ws);
    }
}

void orc_velocity_forcing_field(orc_block *b)
{
    int64_t I = b->ibCount;
    for (int64_t n = 1; n <= I; ++n) {
        int64_t i = LIST(b->intercepted, I, n, 1), j = LIST(b->intercepted, I, n, 2), k = LIST(b->intercepted, I, n, 3);
        lookup_t lk;
        lk_window(&lk, i, j, k);
        int64_t e = V1(b->nelu2, n);
        FLD(b, u, i, j, k) = recon_velocity(b, K_U, e, V1(b->u2NormDis, n), V1(b->xu, i + 1), V1(b->yu, j),
                                            V1(b->zu, k), i, j, k, &lk, 1.0, surf_u(b, e));
        e = V1(b->nelu1, n);
        FLD(b, u, i - 1, j, k) = recon_velocity(b, K_U, e, V1(b->u1NormDis, n), V1(b->xu, i), V1(b->yu, j),
                                                V1(b->zu, k), i, j, k, &lk, 1.0, surf_u(b, e));
        e = V1(b->nelv2, n);
        double vs = surf_v(b, e, 1);
        FLD(b, v, i, j, k) = recon_velocity(b, K_V, e, V1(b->v2NormDis, n), V1(b->xv, i), V1(b->yv, j + 1),
                                            V1(b->zv, k), i, j, k, &lk, 1.0, vs);
        e = V1(b->nelv1, n);
        vs = surf_v(b, e, 1);
        FLD(b, v, i, j - 1, k) = recon_velocity(b, K_V, e, V1(b->v1NormDis, n), V1(b->xv, i), V1(b->yv, j),
                                                V1(b->zv, k), i, j, k, &lk, 1.0, vs);
        e = V1(b->nelw2, n);
        double ws = surf_w(b, e);
        FLD(b, w, i, j, k) = recon_velocity(b, K_W, e, V1(b->w2NormDis, n), V1(b->xw, i), V1(b->yw, j),
                                            V1(b->zw, k + 1), i, j, k, &lk, 1.0, ws);
        e = V1(b->nelw1, n);
        ws = surf_w(b, e);
        FLD(b, w, i, j, k - 1) = recon_velocity(b, K_W, e, V1(b->w1NormDis, n), V1(b->xw, i), V1(b->yw, j),
                                                V1(b->zw, k), i, j, k, &lk, 1.0, ws);
    }
}

void orc_velocity_forcing_ghost(orc_block *b)
{
    int64_t T = b->tsCount;
    for (int64_t n = 1; n <= T; ++n) {
        int64_t i = LIST(b->ts, T, n, 1), j = LIST(b->ts, T, n, 2), k = LIST(b->ts, T, n, 3);
        int64_t t = V1(b->index_ts, n);
        lookup_t lk;
        lk_window(&lk, i, j, k);
        int64_t e = V1(b->nelu2, t);
        V1(b->ghost[G_U2], n) = recon_velocity(b, K_U, e, -V1(b->u2NormDis, t), V1(b->xu, i + 1), V1(b->yu, j),
                                               V1(b->zu, k), i, j, k, &lk, -1.0, surf_u(b, e));
        V1(b->ghost[G_U2T], n) = FLD(b, u, i, j, k);
        e = V1(b->nelu1, t);
        V1(b->ghost[G_U1], n) = recon_velocity(b, K_U, e, -V1(b->u1NormDis, t), V1(b->xu, i), V1(b->yu, j),
                                               V1(b->zu, k), i, j, k, &lk, -1.0, surf_u(b, e));
        V1(b->ghost[G_U1T], n) = FLD(b, u, i - 1, j, k);
        e = V1(b->nelv2, t);
        double vs = (V1(b->surfId, e) == 50) ? 0.0 + b->ydot : surf_v(b, e, 1);
        V1(b->ghost[G_V2], n) = recon_velocity(b, K_V, e, -V1(b->v2NormDis, t), V1(b->xv, i), V1(b->yv, j + 1),
                                               V1(b->zv, k), i, j, k, &lk, -1.0, vs);
        V1(b->ghost[G_V2T], n) = FLD(b, v, i, j, k);
        e = V1(b->nelv1, t);
        vs = surf_v(b, e, 1);
        V1(b->ghost[G_V1], n) = recon_velocity(b, K_V, e, -V1(b->v1NormDis, t), V1(b->xv, i), V1(b->yv, j),
                                               V1(b->zv, k), i, j, k, &lk, -1.0, vs);
        V1(b->ghost[G_V1T], n) = FLD(b, v, i, j - 1, k);
        e = V1(b->nelw2, t);
        double ws = surf_w(b, e);
        V1(b->ghost[G_W2], n) = recon_velocity(b, K_W, e, -V1(b->w2NormDis, t), V1(b->xw, i), V1(b->yw, j),
                                               V1(b->zw, k + 1), i, j, k, &lk, -1.0, ws);
        V1(b->ghost[G_W2T], n) = FLD(b, w, i, j, k);
        e = V1(b->nelw1, t);
        ws = surf_w(b, e);
        V1(b->ghost[G_W1], n) = recon_velocity(b, K_W, e, -V1(b->w1NormDis, t), V1(b->xw, i), V1(b->yw, j),
                                               V1(b->zw, k), i, j, k, &lk, -1.0, ws);
        V1(b->ghost[G_W1T], n) = FLD(b, w, i, j, k - 1);
    }
}

/* ---------------- StressCalc.f90:14-208 ---------------- */
void orc_stress_calculation(orc_block *b, double mu_f, double rho_f)
{
    double viscousDrag = 0.0, pressureDrag = 0.0, viscousLift = 0.0, pressureLift = 0.0;
    double surf_area = 0.0, area_Sx = 0.0, area_Sy = 0.0;
    for (int64_t e = 1; e <= b->ibElems; ++e) {
        double nrm[3] = { V1(b->cosA, e), V1(b->cosB, e), V1(b->cosG, e) };
        int64_t ic = orc_find_index_in_array(V1(b->xcent, e), b->x1, 2, b->nx + 1, NULL);
        int64_t jc = orc_find_index_in_array(V1(b->ycent, e), b->y1, 2, b->ny + 1, NULL);
        int64_t kc = orc_find_index_in_array(V1(b->zcent, e), b->z1, 2, b->nz + 1, NULL);
        double dX = V1(b->x1, ic + 1) - V1(b->x1, ic);
        double dY = V1(b->y1, jc + 1) - V1(b->y1, jc);
        double dZ = V1(b->z1, kc + 1) - V1(b->z1, kc);
        double diag = sqrt(dX * dX + dY * dY + dZ * dZ);
        double nd = diag;
        double px = V1(b->xcent, e) + nd * nrm[0];
        double py = V1(b->ycent, e) + nd * nrm[1];
        double pz = V1(b->zcent, e) + nd * nrm[2];
        int64_t sid = V1(b->surfId, e);
        if (sid == 50) { b->thetaDot = 0.0; b->thetaDDot = 0.0; }
        else if (sid == 51) { b->thetaDot = b->thetaDot1; b->thetaDDot = b->thetaDDot1; }
        else if (sid == 52) { b->thetaDot = b->thetaDot2; b->thetaDDot = b->thetaDDot2; }
        double us = b->xdot;
        double vs = -b->thetaDot * (V1(b->zcent, e) - b->piv_z) + b->ydot;
        double ws = b->thetaDot * (V1(b->ycent, e) - b->piv_y);
        double ac_z = -sq(b->thetaDot) * (V1(b->zcent, e) - b->piv_z);
        double ac_y = -sq(b->thetaDot) * (V1(b->ycent, e) - b->piv_y);
        double at_z = b->thetaDDot * (V1(b->ycent, e) - b->piv_y);
        double at_y = -b->thetaDDot * (V1(b->zcent, e) - b->piv_z);
        double dpdn = -((ac_z + at_z) * V1(b->cosG, e) + (ac_y + at_y) * V1(b->cosB, e)) - b->yddot * V1(b->cosB, e);
        double ddn[3], val, d[3];
        const double *grids[4][3] = { { b->xu, b->yu, b->zu }, { b->xv, b->yv, b->zv },
                                      { b->xw, b->yw, b->zw }, { b->xp, b->yp, b->zp } };
        const double *vars[4] = { b->u, b->v, b->w, b->p };
        double surf[3] = { us, vs, ws };
        for (int c = 0; c < 3; ++c) {
            int64_t ix = orc_find_index_in_array(px, grids[c][0], 2, b->nx + 1, NULL);
            int64_t iy = orc_find_index_in_array(py, grids[c][1], 2, b->ny + 1, NULL);
            int64_t iz = orc_find_index_in_array(pz, grids[c][2], 2, b->nz + 1, NULL);
            value_and_derivs(b, px, py, pz, ix, iy, iz, grids[c][0], grids[c][1], grids[c][2], c + 1,
                             vars[c], &val, d);
            double dn = d[0] * nrm[0] + d[1] * nrm[1] + d[2] * nrm[2];
            ddn[c] = (2.0 / nd) * (val - surf[c]) - dn;
        }
        double alen = V1(b->elen, e);
        double area = alen / 2.0;
        double area_yz = 0.5 * fabs(alen * nrm[0]);
        double area_xz = 0.5 * fabs(alen * nrm[1]);
        double dd = ddn[0] * nrm[0] + ddn[1] * nrm[1] + ddn[2] * nrm[2];
        double shear[3];
        for (int c = 0; c < 3; ++c) shear[c] = (ddn[c] - dd * nrm[c]) * mu_f * area;
        int64_t ix = orc_find_index_in_array(px, b->xp, 2, b->nx + 1, NULL);
        int64_t iy = orc_find_index_in_array(py, b->yp, 2, b->ny + 1, NULL);
        int64_t iz = orc_find_index_in_array(pz, b->zp, 2, b->nz + 1, NULL);
        value_and_derivs(b, px, py, pz, ix, iy, iz, b->xp, b->yp, b->zp, 0, b->p, &val, d);
        double dpdn_e = d[0] * nrm[0] + d[1] * nrm[1] + d[2] * nrm[2];
        double bval = dpdn;
        double aval = (dpdn_e - dpdn) / (2.0 * diag);
        (void)aval;
        (void)bval;
        double cval = val - (dpdn_e + dpdn) * diag * 0.5;
        double fs[3];
        for (int c = 0; c < 3; ++c) fs[c] = -cval * area * nrm[c] * rho_f;
        viscousDrag = viscousDrag + shear[0];
        pressureDrag = pressureDrag + fs[0];
        viscousLift = viscousLift + shear[1];
        pressureLift = pressureLift + fs[1];
        surf_area = surf_area + area;
        area_Sx = area_Sx + area_xz;
        area_Sy = area_Sy + area_yz;
    }
    (void)surf_area;
    area_Sx = 0.5 * area_Sx;
    area_Sy = 0.5 * area_Sy;
    b->viscous_drag = 2.0 * (viscousDrag / area_Sx);
    b->pressure_drag = 2.0 * (pressureDrag / area_Sy);
    b->viscous_lift = 2.0 * (viscousLift / area_Sx);
    b->pressure_lift = 2.0 * (pressureLift / area_Sy);
}

long orc_find_misses(void) { return g_find_miss; }
