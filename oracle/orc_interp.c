/*
 * Test infrastructure: FP64 restatement of the coarse/fine block coupling,
 * reference src/InterfaceDetail.f90 (interfaceDetail :10-165,
 * set_interface_detail :168-241), src/CoarseInterp.f90 (coarseUpdate :16-204,
 * coarseUpdate_newv :206-346, coarseUpdate_pc :347-437,
 * trilinear_interpolation :447-502) and src/FineInterp_bound.f90
 * (fineUpdate_pc_bd :17-107, fineUpdate_newv_bd :109-309,
 * fineUpdate_bd_mv :311-544, trilinear_interpolation :554-603).
 */
#include "orc_internal.h"

enum { C_P = 0, C_U = 1, C_V = 2, C_W = 3 };

#define DET(it, c, a, r, n) ((it)->det[c][a][(r) - 1 + 3 * ((int64_t)(n) - 1)])

static void set_detail(const double *var, int64_t size, double s, double e, int64_t factor,
                       int64_t *det, int64_t *counter, int64_t starter)
{
    int64_t xx1 = 0, xx2 = 0;
    for (int64_t i = 1; i <= size; ++i)
        if (V1(var, i) > s) { xx1 = i; break; }
    for (int64_t i = xx1; i <= size; ++i)
        if (V1(var, i) > e) { xx2 = i - 1; break; }
#define D(r, n) det[(r) - 1 + 3 * ((n) - 1)]
    D(1, 1) = xx1 - 1;
    D(2, 1) = starter - 1;
    D(3, 1) = starter - 1;
    int64_t ender = starter + factor - 1;
    int64_t cnt = 2;
    for (;;) {
        D(1, cnt) = xx1;
        D(2, cnt) = starter;
        D(3, cnt) = ender;
        starter = ender + 1;
        ender = ender + factor;
        cnt = cnt + 1;
        xx1 = xx1 + 1;
        if (xx1 > xx2) break;
    }
    D(1, cnt) = xx2 + 1;
    D(2, cnt) = starter;
    D(3, cnt) = starter;
#undef D
    *counter = cnt;
}

void orc_interface_detail(orc_case *c)
{
    for (int64_t g = 1; g <= c->nintf; ++g) {
        orc_interface *it = &c->intf[g - 1];
        orc_block *a = &c->blk[it->a_blk - 1];
        orc_block *b = &c->blk[it->b_blk - 1];
        int64_t factor = it->b_msh / it->a_msh;
        it->cap[0] = b->nx / factor + 3;
        it->cap[1] = b->ny / factor + 3;
        it->cap[2] = b->nz / factor + 3;
        for (int q = 0; q < 4; ++q)
            for (int ax = 0; ax < 3; ++ax) {
                free(it->det[q][ax]);
                /* one spare column: the table can fill its last entry */
                it->det[q][ax] = orc_xcalloc((size_t)(3 * (it->cap[ax] + 2)), sizeof(int64_t));
            }
        const double *gx[4] = { a->xp, a->xu, a->xv, a->xw };
        const double *gy[4] = { a->yp, a->yu, a->yv, a->yw };
        const double *gz[4] = { a->zp, a->zu, a->zv, a->zw };
        int64_t sx[4] = { a->nx + 2, a->nx + 3, a->nx + 2, a->nx + 2 };
        int64_t sy[4] = { a->ny + 2, a->ny + 2, a->ny + 3, a->ny + 2 };
        int64_t sz[4] = { a->nz + 2, a->nz + 2, a->nz + 2, a->nz + 3 };
        for (int q = 0; q < 4; ++q) {
            set_detail(gx[q], sx[q], it->xs, it->xe, factor, it->det[q][0], &it->counter[q][0], q == C_U ? 3 : 2);
            set_detail(gy[q], sy[q], it->ys, it->ye, factor, it->det[q][1], &it->counter[q][1], q == C_V ? 3 : 2);
            set_detail(gz[q], sz[q], it->zs, it->ze, factor, it->det[q][2], &it->counter[q][2], q == C_W ? 3 : 2);
        }
    }
}

/* CoarseInterp.f90:447-502 */
static double tri_coarse(const orc_block *f, double x, double y, double z, int64_t i, int64_t j,
                         int64_t k, const double *xg, const double *yg, const double *zg,
                         int off, const double *var)
{
    int64_t ii = i, jj = j, kk = k, vi = i, vj = j, vk = k;
    if (off == 1) { ii = i - 1; vi = i - 2; }
    else if (off == 2) { jj = j - 1; vj = j - 2; }
    else if (off == 3) { kk = k - 1; vk = k - 2; }
    double f1[4] = { var[IX(f, vi, vj, vk)], var[IX(f, vi + 1, vj, vk)], var[IX(f, vi, vj + 1, vk)],
                     var[IX(f, vi + 1, vj + 1, vk)] };
    double f2[4] = { var[IX(f, vi, vj, vk + 1)], var[IX(f, vi + 1, vj, vk + 1)], var[IX(f, vi, vj + 1, vk + 1)],
                     var[IX(f, vi + 1, vj + 1, vk + 1)] };
    double z1 = orc_bilinear_interpolation(x, y, V1(xg, ii), V1(xg, ii + 1), V1(yg, jj), V1(yg, jj + 1), f1);
    double z2 = orc_bilinear_interpolation(x, y, V1(xg, ii), V1(xg, ii + 1), V1(yg, jj), V1(yg, jj + 1), f2);
    return orc_linear_interpolation(z, V1(zg, kk), V1(zg, kk + 1), z1, z2);
}

/* FineInterp_bound.f90:554-603 */
static double tri_bound(const orc_block *cb, double x, double y, double z, int64_t i, int64_t j,
                        int64_t k, const double *xg, const double *yg, const double *zg, int off,
                        const double *var)
{
    int64_t vi = i, vj = j, vk = k;
    if (off == 1) vi = i - 1;
    else if (off == 2) vj = j - 1;
    else if (off == 3) vk = k - 1;
    double f1[4] = { var[IX(cb, vi - 1, vj - 1, vk - 1)], var[IX(cb, vi + 1, vj - 1, vk - 1)],
                     var[IX(cb, vi - 1, vj + 1, vk - 1)], var[IX(cb, vi + 1, vj + 1, vk - 1)] };
    double f2[4] = { var[IX(cb, vi - 1, vj - 1, vk + 1)], var[IX(cb, vi + 1, vj - 1, vk + 1)],
                     var[IX(cb, vi - 1, vj + 1, vk + 1)], var[IX(cb, vi + 1, vj + 1, vk + 1)] };
    double z1 = orc_bilinear_interpolation(x, y, V1(xg, i - 1), V1(xg, i + 1), V1(yg, j - 1), V1(yg, j + 1), f1);
    double z2 = orc_bilinear_interpolation(x, y, V1(xg, i - 1), V1(xg, i + 1), V1(yg, j - 1), V1(yg, j + 1), f2);
    return orc_linear_interpolation(z, V1(zg, k - 1), V1(zg, k + 1), z1, z2);
}

static void grids_of(const orc_block *b, int q, const double **gx, const double **gy, const double **gz)
{
    switch (q) {
    case C_P: *gx = b->xp; *gy = b->yp; *gz = b->zp; break;
    case C_U: *gx = b->xu; *gy = b->yu; *gz = b->zu; break;
    case C_V: *gx = b->xv; *gy = b->yv; *gz = b->zv; break;
    default: *gx = b->xw; *gy = b->yw; *gz = b->zw; break;
    }
}

/* fine -> coarse restriction of component q from fvar (fine) into cvar (coarse);
 * cvar2 (if not NULL) receives the same value (pc and pco) */
static void coarse_comp(orc_case *c, orc_interface *it, int q, const double *fvar_of_b,
                        double *cvar, double *cvar2)
{
    orc_block *a = &c->blk[it->a_blk - 1];
    orc_block *b = &c->blk[it->b_blk - 1];
    const double *cx, *cy, *cz, *fx, *fy, *fz;
    grids_of(a, q, &cx, &cy, &cz);
    grids_of(b, q, &fx, &fy, &fz);
    int64_t ci = b->cintp;
    for (int64_t k = ci; k <= it->counter[q][2] - ci; ++k)
        for (int64_t j = ci; j <= it->counter[q][1] - ci; ++j)
            for (int64_t i = ci; i <= it->counter[q][0] - ci; ++i) {
                int64_t tx = DET(it, q, 0, 1, i), ty = DET(it, q, 1, 1, j), tz = DET(it, q, 2, 1, k);
                int64_t lx = (DET(it, q, 0, 2, i) + DET(it, q, 0, 3, i) - 1) / 2;
                int64_t ly = (DET(it, q, 1, 2, j) + DET(it, q, 1, 3, j) - 1) / 2;
                int64_t lz = (DET(it, q, 2, 2, k) + DET(it, q, 2, 3, k) - 1) / 2;
                int64_t ox = 0, oy = 0, oz = 0;
                if (q == C_U) { lx = DET(it, q, 0, 3, i); ox = 1; }
                if (q == C_V) { ly = DET(it, q, 1, 3, j); oy = 1; }
                if (q == C_W) { lz = DET(it, q, 2, 3, k); oz = 1; }
                double val = tri_coarse(b, V1(cx, tx), V1(cy, ty), V1(cz, tz), lx, ly, lz, fx, fy, fz, q, fvar_of_b);
                cvar[IX(a, tx - ox, ty - oy, tz - oz)] = val;
                if (cvar2) cvar2[IX(a, tx - ox, ty - oy, tz - oz)] = val;
            }
}

void orc_coarse_update(orc_case *c)
{
    for (int64_t g = 1; g <= c->nintf; ++g) {
        orc_interface *it = &c->intf[g - 1];
        orc_block *a = &c->blk[it->a_blk - 1], *b = &c->blk[it->b_blk - 1];
        coarse_comp(c, it, C_P, b->p, a->p, NULL);
        coarse_comp(c, it, C_U, b->u, a->u, NULL);
        coarse_comp(c, it, C_V, b->v, a->v, NULL);
        coarse_comp(c, it, C_W, b->w, a->w, NULL);
    }
}

void orc_coarse_update_newv(orc_case *c)
{
    for (int64_t g = 1; g <= c->nintf; ++g) {
        orc_interface *it = &c->intf[g - 1];
        orc_block *a = &c->blk[it->a_blk - 1], *b = &c->blk[it->b_blk - 1];
        coarse_comp(c, it, C_U, b->ut, a->ut, NULL);
        coarse_comp(c, it, C_V, b->vt, a->vt, NULL);
        coarse_comp(c, it, C_W, b->wt, a->wt, NULL);
    }
}

void orc_coarse_update_pc(orc_case *c)
{
    for (int64_t g = 1; g <= c->nintf; ++g) {
        orc_interface *it = &c->intf[g - 1];
        orc_block *a = &c->blk[it->a_blk - 1], *b = &c->blk[it->b_blk - 1];
        coarse_comp(c, it, C_P, b->pc, a->pc, a->pco);
    }
}

/* coarse -> fine on the six faces of the fine block, component q */
static void fine_comp(orc_case *c, orc_interface *it, int q, const double *cvar, double *fvar,
                      double *fvar2)
{
    orc_block *a = &c->blk[it->a_blk - 1];
    orc_block *b = &c->blk[it->b_blk - 1];
    const double *cx, *cy, *cz, *fx, *fy, *fz;
    grids_of(a, q, &cx, &cy, &cz);
    grids_of(b, q, &fx, &fy, &fz);
    int64_t ox = (q == C_U), oy = (q == C_V), oz = (q == C_W);
    for (int axis = 1; axis <= 3; ++axis) {
        int64_t st[3] = { 1, 1, 1 };
        st[axis - 1] = it->counter[q][axis - 1] - 1;
        for (int64_t k = 1; k <= it->counter[q][2]; k += st[2])
            for (int64_t j = 1; j <= it->counter[q][1]; j += st[1])
                for (int64_t i = 1; i <= it->counter[q][0]; i += st[0]) {
                    int64_t x1 = DET(it, q, 0, 2, i), x2 = DET(it, q, 0, 3, i);
                    int64_t y1 = DET(it, q, 1, 2, j), y2 = DET(it, q, 1, 3, j);
                    int64_t z1 = DET(it, q, 2, 2, k), z2 = DET(it, q, 2, 3, k);
                    int64_t lx = DET(it, q, 0, 1, i), ly = DET(it, q, 1, 1, j), lz = DET(it, q, 2, 1, k);
                    for (int64_t tz = z1; tz <= z2; ++tz)
                        for (int64_t ty = y1; ty <= y2; ++ty)
                            for (int64_t tx = x1; tx <= x2; ++tx) {
                                double val = tri_bound(a, V1(fx, tx), V1(fy, ty), V1(fz, tz), lx, ly, lz,
                                                       cx, cy, cz, q, cvar);
                                fvar[IX(b, tx - ox, ty - oy, tz - oz)] = val;
                                if (fvar2) fvar2[IX(b, tx - ox, ty - oy, tz - oz)] = val;
                            }
                }
    }
}

void orc_fine_update_pc_bd(orc_case *c)
{
    for (int64_t g = 1; g <= c->nintf; ++g) {
        orc_interface *it = &c->intf[g - 1];
        orc_block *a = &c->blk[it->a_blk - 1], *b = &c->blk[it->b_blk - 1];
        fine_comp(c, it, C_P, a->pc, b->pc, b->pco);
    }
}

void orc_fine_update_newv_bd(orc_case *c)
{
    for (int64_t g = 1; g <= c->nintf; ++g) {
        orc_interface *it = &c->intf[g - 1];
        orc_block *a = &c->blk[it->a_blk - 1], *b = &c->blk[it->b_blk - 1];
        fine_comp(c, it, C_U, a->ut, b->ut, NULL);
        fine_comp(c, it, C_V, a->vt, b->vt, NULL);
        fine_comp(c, it, C_W, a->wt, b->wt, NULL);
    }
}

void orc_fine_update_bd_mv(orc_case *c, int64_t g)
{
    orc_interface *it = &c->intf[g - 1];
    orc_block *a = &c->blk[it->a_blk - 1], *b = &c->blk[it->b_blk - 1];
    fine_comp(c, it, C_P, a->p, b->p, NULL);
    fine_comp(c, it, C_U, a->u, b->u, NULL);
    fine_comp(c, it, C_V, a->v, b->v, NULL);
    fine_comp(c, it, C_W, a->w, b->w, NULL);
}
