/*
 * Test infrastructure: FP64 restatement of the reference's block-move path,
 * src/Search.F90 (cellCount_solid_coarse_mv :916-995, block_move_check
 * :997-1203, change_block_coords :1208-1264, change_block_coords_interfaces
 * :1269-1370, change_block_interface :1372-1461, tagging_th_move :415-432)
 * and src/FineInterp.f90 (fineUpdate_mv :10-281), called in the order of
 * app/main.F90:171-229.
 *
 * The reference's block_move_check reads the marker nodes mk and mkx1 that
 * findDistnode sets only from ID-51 nodes (Search.F90:21-47, :1018-1021);
 * the driver runs it only when every fine block has them
 * (orc_move_markers_defined).  A copy range that would read outside the
 * block (undefined in the reference) stops the oracle with a message.
 */
#include <stdio.h>
#include <string.h>

#include "orc_internal.h"

#define DET(it, c, a, r, n) ((it)->det[c][a][(r) - 1 + 3 * ((int64_t)(n) - 1)])

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

int orc_move_markers_defined(const orc_case *c)
{
    for (int64_t g = 1; g <= c->nintf; ++g) {
        const orc_block *b = &c->blk[c->intf[g - 1].b_blk - 1];
        if (b->mk < 1 || b->mkx1 < 1 || b->mkx2 < 1) return 0;
    }
    return 1;
}

/* a shift of m cells rounded to whole coarse cells, at least one
 * (Search.F90:1043-1059) */
static int64_t whole_coarse(int64_t m, int64_t factor)
{
    int64_t am = m < 0 ? -m : m;
    if (am < factor) return m < 0 ? -factor : factor;
    return m < 0 ? m + am % factor : m - am % factor;
}

/* first centre above lo, and the last centre not above hi after it; a bound
 * with no such centre keeps its previous value (Search.F90:1129-1184) */
static void copy_range(const double *c, int64_t n2, double lo, double hi, int64_t *start, int64_t *end)
{
    for (int64_t i = 1; i <= n2; ++i)
        if (V1(c, i) > lo) { *start = i; break; }
    for (int64_t i = *start < 1 ? 1 : *start; i <= n2; ++i)
        if (V1(c, i) > hi) { *end = i - 1; break; }
}

void orc_block_move_check(orc_case *c, int64_t *coarse_flcnt_check)
{
    for (int64_t g = 1; g <= c->nintf; ++g) {
        orc_interface *it = &c->intf[g - 1];
        orc_block *b = &c->blk[it->b_blk - 1];
        if (!b->u) continue;
        double mg1 = (double)b->cintp * c->blk[0].dx;
        double marginx = 4.5 * mg1, marginy = 5.0 * mg1, marginz = 1.52 * mg1;
        int64_t factor = it->b_msh / it->a_msh;
        double ymk = V1(b->ynode1, b->mk);
        double yval_up = dmax(ymk, b->nxty_cent + 0.02);
        double yval_dw = dmin(ymk, b->nxty_cent - 0.02);
        double xval_lt = V1(b->xnode1, b->mkx1);
        b->move_amty = 0;
        b->move_amtx = 0;
        b->move_amtz = 0;
        if (!(fabs(xval_lt - it->xs) <= marginx || fabs(b->piv_z - it->zs) <= marginz ||
              fabs(b->piv_z - it->ze) <= marginz || fabs(yval_dw - it->ys) <= marginy ||
              fabs(yval_up - it->ye) <= marginy))
            continue;
        b->move_check = 1;
        b->blk_mv_tag = 1;
        *coarse_flcnt_check = 1;
        if (fabs(yval_dw - it->ys) <= marginy || fabs(yval_up - it->ye) <= marginy) {
            b->move_amty = whole_coarse((int64_t)floor((b->nxty_cent - b->inity_cent) / b->dy), factor);
            b->inity_cent = b->nxty_cent;
        }
        if (fabs(xval_lt - it->xs) < marginx) {
            b->move_amtx = whole_coarse((int64_t)floor((b->nxtx_cent - b->initx_cent) / b->dx) - factor, factor);
            b->initx_cent = b->nxtx_cent;
        }
        if (fabs(b->piv_z - it->zs) < (0.500 - 0.410 - mg1) || fabs(b->piv_z - it->ze) < (0.600 - 0.500 - mg1)) {
            b->move_amtz = whole_coarse((int64_t)floor((b->nxtz_cent - b->initz_cent) / b->dz), factor);
            b->initz_cent = b->nxtz_cent;
        }
        double sx = (double)b->move_amtx * b->dx, sy = (double)b->move_amty * b->dy,
               sz = (double)b->move_amtz * b->dz;
        it->xs_new = dmax(it->xs, it->xs + (double)b->move_amtx * b->dx);
        it->xe_new = dmin(it->xe, it->xe + (double)b->move_amtx * b->dx);
        it->ys_new = dmax(it->ys, it->ys + (double)b->move_amty * b->dy);
        it->ye_new = dmin(it->ye, it->ye + (double)b->move_amty * b->dy);
        it->zs_new = dmax(it->zs, it->zs + (double)b->move_amtz * b->dz);
        it->ze_new = dmin(it->ze, it->ze + (double)b->move_amtz * b->dz);
        it->xs = it->xs + sx;
        it->xe = it->xe + sx;
        it->ys = it->ys + sy;
        it->ye = it->ye + sy;
        it->zs = it->zs + sz;
        it->ze = it->ze + sz;
        copy_range(b->xp, NX2(b), it->xs_new, it->xe_new, &b->cpy[0], &b->cpy[1]);
        copy_range(b->yp, NY2(b), it->ys_new, it->ye_new, &b->cpy[2], &b->cpy[3]);
        copy_range(b->zp, NZ2(b), it->zs_new, it->ze_new, &b->cpy[4], &b->cpy[5]);
        size_t n2 = (size_t)(NX2(b) * NY2(b) * NZ2(b));
        if (!b->u_dum) {
            b->u_dum = orc_xcalloc(n2, sizeof(double));
            b->v_dum = orc_xcalloc(n2, sizeof(double));
            b->w_dum = orc_xcalloc(n2, sizeof(double));
            b->p_dum = orc_xcalloc(n2, sizeof(double));
        }
        memcpy(b->u_dum, b->u, n2 * sizeof(double));
        memcpy(b->v_dum, b->v, n2 * sizeof(double));
        memcpy(b->w_dum, b->w, n2 * sizeof(double));
        memcpy(b->p_dum, b->p, n2 * sizeof(double));
    }
}

void orc_change_block_coords(orc_block *b)
{
    if (b->move_check != 1) return;
    double cz = (double)b->move_amtz * b->dz;
    double cy = (double)b->move_amty * b->dy;
    double cx = (double)b->move_amtx * b->dx;
    for (int64_t i = 1; i <= b->nx + 3; ++i) V1(b->x1, i) = V1(b->x1, i) + cx;
    for (int64_t i = 1; i <= b->ny + 3; ++i) V1(b->y1, i) = V1(b->y1, i) + cy;
    for (int64_t i = 1; i <= b->nz + 3; ++i) V1(b->z1, i) = V1(b->z1, i) + cz;
    for (int64_t i = 1; i <= b->nx + 3; ++i) V1(b->xu, i) = V1(b->x1, i);
    for (int64_t i = 1; i <= b->ny + 3; ++i) V1(b->yv, i) = V1(b->y1, i);
    for (int64_t i = 1; i <= b->nz + 3; ++i) V1(b->zw, i) = V1(b->z1, i);
    for (int64_t i = 1; i <= b->ny + 2; ++i) {
        V1(b->yu, i) = 0.5 * (V1(b->y1, i) + V1(b->y1, i + 1));
        V1(b->yw, i) = V1(b->yu, i);
        V1(b->yp, i) = V1(b->yu, i);
    }
    for (int64_t i = 1; i <= b->nx + 2; ++i) {
        V1(b->xv, i) = 0.5 * (V1(b->x1, i) + V1(b->x1, i + 1));
        V1(b->xw, i) = V1(b->xv, i);
        V1(b->xp, i) = V1(b->xv, i);
    }
    for (int64_t i = 1; i <= b->nz + 2; ++i) {
        V1(b->zu, i) = 0.5 * (V1(b->z1, i) + V1(b->z1, i + 1));
        V1(b->zv, i) = V1(b->zu, i);
        V1(b->zp, i) = V1(b->zu, i);
    }
}

static void range_check(const char *what, int64_t lo, int64_t hi, int64_t n)
{
    if (lo < 1 || hi > n) {
        fprintf(stderr, "oracle: block move copies %s indices %lld..%lld outside 1..%lld\n", what, (long long)lo,
                (long long)hi, (long long)n);
        exit(4);
    }
}

void orc_change_block_coords_interfaces(orc_interface *it, orc_block *b)
{
    if (!b->u) return;
    if (b->move_check != 1) return;
    for (int64_t k = 1; k <= NZ2(b); ++k)
        for (int64_t j = 1; j <= NY2(b); ++j)
            for (int64_t i = 1; i <= NX2(b); ++i) {
                double xp = V1(b->xp, i), yp = V1(b->yp, j), zp = V1(b->zp, k);
                FLD(b, cell_n, i, j, k) =
                    (xp > (it->xs_new + b->dx) && xp < (it->xe_new - b->dx) && zp > (it->zs_new + b->dz) &&
                     zp < (it->ze_new - b->dz) && yp > (it->ys_new + b->dy) && yp < (it->ye_new - b->dy))
                        ? 1 : 0;
            }
    copy_range(b->xp, NX2(b), it->xs_new, it->xe_new, &b->cpy_mv[0], &b->cpy_mv[1]);
    copy_range(b->yp, NY2(b), it->ys_new, it->ye_new, &b->cpy_mv[2], &b->cpy_mv[3]);
    copy_range(b->zp, NZ2(b), it->zs_new, it->ze_new, &b->cpy_mv[4], &b->cpy_mv[5]);
    size_t n2 = (size_t)(NX2(b) * NY2(b) * NZ2(b));
    memset(b->u, 0, n2 * sizeof(double));
    memset(b->v, 0, n2 * sizeof(double));
    memset(b->w, 0, n2 * sizeof(double));
    memset(b->p, 0, n2 * sizeof(double));
    const int64_t *m = b->cpy_mv, *o = b->cpy;
    if (m[1] < m[0] || m[3] < m[2] || m[5] < m[4]) return;
    range_check("target x", m[0], m[1], NX2(b));
    range_check("target y", m[2], m[3], NY2(b));
    range_check("target z", m[4], m[5], NZ2(b));
    range_check("source x", o[0], o[0] + (m[1] - m[0]), NX2(b));
    range_check("source y", o[2], o[2] + (m[3] - m[2]), NY2(b));
    range_check("source z", o[4], o[4] + (m[5] - m[4]), NZ2(b));
    for (int64_t k = m[4], sk = o[4]; k <= m[5]; ++k, ++sk)
        for (int64_t i = m[0], si = o[0]; i <= m[1]; ++i, ++si)
            for (int64_t j = m[2], sj = o[2]; j <= m[3]; ++j, ++sj) {
                int64_t d = IX(b, i, j, k), s = IX(b, si, sj, sk);
                b->u[d] = b->u_dum[s];
                b->v[d] = b->v_dum[s];
                b->w[d] = b->w_dum[s];
                b->p[d] = b->p_dum[s];
            }
}

void orc_change_block_interface(orc_interface *it, const orc_block *b)
{
    if (!b->u) return;
    if (b->move_check != 1) return;
    int64_t factor = it->b_msh / it->a_msh;
    const int64_t inc[3] = { b->move_amtx / factor, b->move_amty / factor, b->move_amtz / factor };
    for (int q = 0; q < 4; ++q)
        for (int a = 0; a < 3; ++a)
            for (int64_t n = 1; n <= it->counter[q][a]; ++n) DET(it, q, a, 1, n) += inc[a];
}

void orc_cell_count_solid_coarse_mv(orc_case *c, int64_t *coarse_flcnt_check)
{
    if (*coarse_flcnt_check == 1) orc_cell_count_solid_coarse(&c->blk[0]);
    *coarse_flcnt_check = 0;
}

void orc_tagging_move(orc_case *c, int64_t g)
{
    orc_block *b = &c->blk[g - 1];
    if (b->move_check != 1) return;
    orc_tagging_core(b);
    for (int64_t q = 1; q <= c->nintf; ++q) {
        if (c->intf[q - 1].b_blk != g) continue;
        orc_fine_update_mv(c, q);
        orc_fine_update_bd_mv(c, q);
    }
}

/* ---------------------------- fineUpdate_mv ---------------------------- */
enum { C_P = 0, C_U = 1, C_V = 2, C_W = 3 };

/* bilinear value in x, y from the four corners, then the linear blend in z
 * of the two planes (FineInterp.f90:50-81) */
static double plane_blend(double vx, double vy, double vz, double x1, double x2, double y1, double y2, double z1,
                          double z2, const double f_lo[4], const double f_hi[4])
{
    double deno = (x2 - x1) * (y2 - y1);
    double xtx = (x2 - vx), xxo = (vx - x1), yty = (y2 - vy), yyo = (vy - y1);
    double first = (f_lo[0] * xtx + f_lo[1] * xxo) * yty;
    double second = (f_lo[3] * xtx + f_lo[2] * xxo) * yyo;
    double ans1 = (first + second) / deno;
    first = (f_hi[0] * xtx + f_hi[1] * xxo) * yty;
    second = (f_hi[3] * xtx + f_hi[2] * xxo) * yyo;
    double ans2 = (first + second) / deno;
    return ans1 + ((vz - z1) * ((ans2 - ans1) / (z2 - z1)));
}

void orc_fine_update_mv(orc_case *c, int64_t g)
{
    orc_interface *it = &c->intf[g - 1];
    orc_block *a = &c->blk[it->a_blk - 1];
    orc_block *b = &c->blk[it->b_blk - 1];
    for (int comp = C_P; comp <= C_W; ++comp) {
        const double *fx, *fy, *fz, *cx, *cy, *cz;
        const double *cv;
        double *fv;
        switch (comp) {
        case C_P: fx = b->xp; fy = b->yp; fz = b->zp; cx = a->xp; cy = a->yp; cz = a->zp; cv = a->p; fv = b->p; break;
        case C_U: fx = b->xu; fy = b->yu; fz = b->zu; cx = a->xu; cy = a->yu; cz = a->zu; cv = a->u; fv = b->u; break;
        case C_V: fx = b->xv; fy = b->yv; fz = b->zv; cx = a->xv; cy = a->yv; cz = a->zv; cv = a->v; fv = b->v; break;
        default:  fx = b->xw; fy = b->yw; fz = b->zw; cx = a->xw; cy = a->yw; cz = a->zw; cv = a->w; fv = b->w; break;
        }
        /* the staggered component's coarse values sit one index lower
         * (u(i) at xu(i+1)), and so does its fine target */
        int ox = comp == C_U, oy = comp == C_V, oz = comp == C_W;
        for (int64_t k = 2; k <= it->counter[comp][2] - 1; ++k)
            for (int64_t j = 2; j <= it->counter[comp][1] - 1; ++j)
                for (int64_t i = 2; i <= it->counter[comp][0] - 1; ++i) {
                    int64_t lx = DET(it, comp, 0, 1, i), ly = DET(it, comp, 1, 1, j), lz = DET(it, comp, 2, 1, k);
                    for (int64_t q = DET(it, comp, 2, 2, k); q <= DET(it, comp, 2, 3, k); ++q)
                        for (int64_t l = DET(it, comp, 1, 2, j); l <= DET(it, comp, 1, 3, j); ++l)
                            for (int64_t s = DET(it, comp, 0, 2, i); s <= DET(it, comp, 0, 3, i); ++s) {
                                if (FLD(b, cell_n, s, l, q) == 1) continue;
                                int64_t i0 = lx - 1 - ox, i1 = lx + 1 - ox;
                                int64_t j0 = ly - 1 - oy, j1 = ly + 1 - oy;
                                int64_t k0 = lz - 1 - oz, k1 = lz + 1 - oz;
                                if (i0 < 1 || j0 < 1 || k0 < 1 || i1 > NX2(a) || j1 > NY2(a) || k1 > NZ2(a) ||
                                    lx - 1 < 1 || ly - 1 < 1 || lz - 1 < 1 || s - ox < 1 || l - oy < 1 ||
                                    q - oz < 1) {
                                    fprintf(stderr, "oracle: fineUpdate_mv stencil outside the blocks\n");
                                    exit(4);
                                }
                                double f_lo[4] = { cv[IX(a, i0, j0, k0)], cv[IX(a, i1, j0, k0)],
                                                   cv[IX(a, i1, j1, k0)], cv[IX(a, i0, j1, k0)] };
                                double f_hi[4] = { cv[IX(a, i0, j0, k1)], cv[IX(a, i1, j0, k1)],
                                                   cv[IX(a, i1, j1, k1)], cv[IX(a, i0, j1, k1)] };
                                double val = plane_blend(V1(fx, s), V1(fy, l), V1(fz, q), V1(cx, lx - 1),
                                                         V1(cx, lx + 1), V1(cy, ly - 1), V1(cy, ly + 1),
                                                         V1(cz, lz - 1), V1(cz, lz + 1), f_lo, f_hi);
                                fv[IX(b, s - ox, l - oy, q - oz)] = val;
                            }
                }
    }
}
