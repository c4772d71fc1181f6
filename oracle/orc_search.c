/*
 * Test infrastructure: FP64 restatement of the immersed-boundary geometry,
 * tagging and list compaction, reference src/Search.F90 (findDistnode :21-47,
 * shiftSurfaceNodesInitial :49-124, computeSurfaceVariables :126-208,
 * computeSurfaceNorm :210-258, tagging_th_core :260-386, findTScells
 * :434-518, selectiveRetagging_th :520-651, cellCount_solid :653-741,
 * computeNormDistance :743-861, fine_block_cell :863-914,
 * cellCount_solid_coarse :1463-1532).  Lists are built in the serial CPU
 * order (i fastest, then j, then k).
 */
#include <string.h>
#include "orc_internal.h"

void orc_find_dist_node(orc_block *b)
{
    double dist = 0.0, dist1 = 999999.0, dist2 = 0.0;
    for (int64_t i = 1; i <= b->ibNodes; ++i) {
        if (V1(b->ibNodeId, i) == 51) {
            if (fabs(V1(b->znode, i)) > dist) { dist = V1(b->znode, i); b->mk = i; }
        }
        if (V1(b->ibNodeId, i) == 51) {
            if (fabs(V1(b->xnode, i)) < dist1) { dist1 = V1(b->xnode, i); b->mkx1 = i; }
            if (fabs(V1(b->xnode, i)) > dist2) { dist2 = V1(b->xnode, i); b->mkx2 = i; }
        }
    }
}

static void rotate_node(const orc_block *b, int64_t i, double aoa1, double aoa2, double piv_pt,
                        double *xr, double *yr, double *zr)
{
    double xn = V1(b->xnode, i), yn = V1(b->ynode, i), zn = V1(b->znode, i);
    if (V1(b->ibNodeId, i) == 51) {
        *xr = xn;
        *zr = zn * cos(aoa1) + yn * sin(aoa1) + piv_pt - piv_pt * cos(aoa1);
        *yr = -zn * sin(aoa1) + yn * cos(aoa1) + piv_pt * sin(aoa1);
    } else if (V1(b->ibNodeId, i) == 52) {
        *xr = xn;
        *zr = zn * cos(aoa2) + yn * sin(aoa2) + piv_pt - piv_pt * cos(aoa2);
        *yr = -zn * sin(aoa2) + yn * cos(aoa2) + piv_pt * sin(aoa2);
    } else {
        *xr = xn; *zr = zn; *yr = yn;
    }
}

void orc_shift_surface_nodes_initial(orc_block *b, double aoa, double piv_pt, int64_t ita,
                                     double dt, double dxmin, double totime)
{
    const double xfact = 0.05;
    double aoa1 = aoa * ORC_PI / 180.0;
    double aoa2 = -aoa1;
    free(b->xnode1); free(b->ynode1); free(b->znode1);
    b->xnode1 = orc_xcalloc((size_t)b->ibNodes, sizeof(double));
    b->ynode1 = orc_xcalloc((size_t)b->ibNodes, sizeof(double));
    b->znode1 = orc_xcalloc((size_t)b->ibNodes, sizeof(double));
    memcpy(b->xnode1, b->xnode, sizeof(double) * (size_t)b->ibNodes);
    memcpy(b->ynode1, b->ynode, sizeof(double) * (size_t)b->ibNodes);
    memcpy(b->znode1, b->znode, sizeof(double) * (size_t)b->ibNodes);
    double bdfr = 15.0;
    double bdy = 15.0 * dxmin;
    double angt = 2.0 * ORC_PI * bdfr;
    b->xmove = 0.0; b->ymove = 0.0; b->zmove = 0.0;
    b->a0 = b->a0 * ORC_PI / 180.0;
    angt = 2.0 * ORC_PI * b->bfreq;
    b->xpth1 = b->xshift - ((double)ita * dxmin * xfact);
    b->xpth2 = b->xshift - ((double)ita * dxmin * xfact);
    b->ypth2 = (b->yamp) * sin(angt * b->xshift);
    b->piv_x = b->xshift; b->piv_y = b->yshift; b->piv_z = b->zshift;
    b->thetaDot = 0.0; b->thetaDDot = 0.0; b->thetaDot1 = 0.0; b->thetaDot2 = 0.0;
    b->alphaDot = 0.0; b->thetaDDot1 = 0.0; b->thetaDDot2 = 0.0;
    b->yt = bdy * sin(2.0 * ORC_PI * bdfr * totime);
    b->ydot = angt * bdy * cos(2.0 * ORC_PI * bdfr * totime);
    b->yddot = -angt * angt * bdy * sin(2.0 * ORC_PI * bdfr * totime);
    b->xt = b->xshift - ((double)ita * dxmin * xfact);
    b->xdot = -(dxmin * xfact) / dt;
    b->inity_cent = b->yshift; b->nxty_cent = b->yshift;
    b->initx_cent = b->xshift; b->nxtx_cent = b->xshift;
    b->initz_cent = b->zshift; b->nxtz_cent = b->zshift;
    for (int64_t i = 1; i <= b->ibNodes; ++i) {
        double xr, yr, zr;
        rotate_node(b, i, aoa1, aoa2, piv_pt, &xr, &yr, &zr);
        V1(b->xnode1, i) = xr + b->xshift;
        V1(b->ynode1, i) = yr + b->yshift;
        V1(b->znode1, i) = zr + b->zshift;
    }
}

void orc_compute_surface_variables(orc_block *b, double phase_angle, double piv_pt,
                                   int64_t ita, double dt, double dxmin, double freq,
                                   double totime)
{
    const double xfact = 0.05;
    double aoa1 = (b->a0) * sin(2.0 * ORC_PI * freq * (totime + dt) + phase_angle);
    double aoa2 = -aoa1;
    double ang_theta = 2.0 * ORC_PI * freq;
    double bdfr = b->bfreq;
    double bdy = b->yamp;
    double angt = 2.0 * ORC_PI * bdfr;
    b->xpth2 = b->xshift - ((double)ita * dxmin * xfact);
    b->ypth2 = (b->yamp) * sin(angt * b->xpth2);
    b->xchg = b->xpth2 - b->xpth1;
    b->xpth1 = b->xpth2;
    b->thetaDot1 = ang_theta * b->a0 * cos(2.0 * ORC_PI * freq * (totime + dt) + phase_angle);
    b->thetaDDot1 = -ang_theta * ang_theta * b->a0 * sin(2.0 * ORC_PI * freq * (totime + dt) + phase_angle);
    b->thetaDot2 = -b->thetaDot1;
    b->thetaDDot2 = -b->thetaDDot1;
    b->yt = bdy * sin(angt * (totime));
    b->ydot = angt * bdy * cos(angt * (totime));
    b->yddot = -angt * angt * bdy * sin(angt * (totime));
    b->xt = b->xpth2;
    b->xdot = -(dxmin * xfact) / dt;
    b->ychg = b->yt - bdy * sin(angt * (totime - dt));
    b->ymove = b->yt;
    b->xmove = b->xchg;
    b->zmove = 0.0;
    b->piv_y = b->piv_y + b->ychg;
    b->piv_x = b->piv_x + b->xmove;
    b->piv_z = b->piv_z + b->zmove;
    b->nxty_cent = b->nxty_cent + b->ychg;
    b->nxtx_cent = b->nxtx_cent + b->xmove;
    for (int64_t i = 1; i <= b->ibNodes; ++i) {
        double xr, yr, zr;
        rotate_node(b, i, aoa1, aoa2, piv_pt, &xr, &yr, &zr);
        V1(b->xnode1, i) = xr + b->xpth2;
        V1(b->ynode1, i) = yr + b->yshift + b->ymove;
        V1(b->znode1, i) = zr + b->zshift + b->zmove;
    }
}

void orc_compute_surface_norm(orc_block *b, int64_t inor)
{
    double dn = (double)inor;
    for (int64_t n = 1; n <= b->ibElems; ++n) {
        int64_t a = V1(b->elP1, n), c = V1(b->elP2, n), d = V1(b->elP3, n);
        double p1x = V1(b->xnode1, a), p1y = V1(b->ynode1, a), p1z = V1(b->znode1, a);
        double p2x = V1(b->xnode1, c), p2y = V1(b->ynode1, c), p2z = V1(b->znode1, c);
        double p3x = V1(b->xnode1, d), p3y = V1(b->ynode1, d), p3z = V1(b->znode1, d);
        V1(b->xcent, n) = (p2x + p1x + p3x) / 3.0;
        V1(b->ycent, n) = (p2y + p1y + p3y) / 3.0;
        V1(b->zcent, n) = (p2z + p1z + p3z) / 3.0;
        double ca = (p2y - p1y) * (p3z - p1z) - (p3y - p1y) * (p2z - p1z);
        double cb = (p2z - p1z) * (p3x - p1x) - (p3z - p1z) * (p2x - p1x);
        double cg = (p2x - p1x) * (p3y - p1y) - (p3x - p1x) * (p2y - p1y);
        double len = sqrt(ca * ca + cb * cb + cg * cg);
        V1(b->elen, n) = len;
        V1(b->cosA, n) = dn * ca / len;
        V1(b->cosB, n) = dn * cb / len;
        V1(b->cosG, n) = dn * cg / len;
    }
}

/* nearest-element pair search for one cell (Search.F90:294-338) */
static void tag_cell(orc_block *b, int64_t i, int64_t j, int64_t k)
{
    double minDis = 1e14, minDis1 = 1e14;
    int64_t nel2Cen = 0, nel2Pnt = 0;
    double xpi = V1(b->xp, i), ypj = V1(b->yp, j), zpk = V1(b->zp, k);
    double x1i = V1(b->x1, i), y1j = V1(b->y1, j), z1k = V1(b->z1, k);
    for (int64_t m = 1; m <= b->ibElems; ++m) {
        double dc = sq(xpi - V1(b->xcent, m)) + sq(ypj - V1(b->ycent, m)) + sq(zpk - V1(b->zcent, m));
        double dp = sq(x1i - V1(b->xcent, m)) + sq(y1j - V1(b->ycent, m)) + sq(z1k - V1(b->zcent, m));
        if (dc < minDis) { minDis = dc; nel2Cen = m; }
        if (dp < minDis1) { minDis1 = dp; nel2Pnt = m; }
    }
    if ((x1i <= V1(b->xcent, nel2Cen) && V1(b->x1, i + 1) >= V1(b->xcent, nel2Cen)) &&
        (y1j <= V1(b->ycent, nel2Cen) && V1(b->y1, j + 1) >= V1(b->ycent, nel2Cen)) &&
        (z1k <= V1(b->zcent, nel2Cen) && V1(b->z1, k + 1) >= V1(b->zcent, nel2Cen)))
        FLD(b, cell, i, j, k) = 2;
    double n2dotn = (x1i - V1(b->xcent, nel2Pnt)) * V1(b->cosA, nel2Pnt) +
                    (y1j - V1(b->ycent, nel2Pnt)) * V1(b->cosB, nel2Pnt) +
                    (z1k - V1(b->zcent, nel2Pnt)) * V1(b->cosG, nel2Pnt);
    TAG3(b, nodeIdTag, i, j, k) = (n2dotn >= -1e-16) ? 0 : 1;
}

static int64_t corner_sum(const orc_block *b, int64_t i, int64_t j, int64_t k)
{
    return TAG3(b, nodeIdTag, i, j, k) + TAG3(b, nodeIdTag, i + 1, j, k) +
           TAG3(b, nodeIdTag, i, j + 1, k) + TAG3(b, nodeIdTag, i + 1, j + 1, k) +
           TAG3(b, nodeIdTag, i, j, k + 1) + TAG3(b, nodeIdTag, i + 1, j, k + 1) +
           TAG3(b, nodeIdTag, i, j + 1, k + 1) + TAG3(b, nodeIdTag, i + 1, j + 1, k + 1);
}

static void count_tags(orc_block *b)
{
    b->ibCount = b->solidCount = b->fluidCount = 0;
    for (int64_t k = 2; k <= b->nz + 1; ++k)
        for (int64_t j = 2; j <= b->ny + 1; ++j)
            for (int64_t i = 2; i <= b->nx + 1; ++i) {
                int64_t c = FLD(b, cell, i, j, k);
                if (c == 1) b->solidCount++;
                else if (c == 0) b->fluidCount++;
                else if (c == 2) b->ibCount++;
            }
}

void orc_tagging_core(orc_block *b)
{
    int64_t n2 = NX2(b) * NY2(b) * NZ2(b);
    int64_t n3 = (b->nx + 3) * (b->ny + 3) * (b->nz + 3);
    memset(b->cell, 0, sizeof(int64_t) * (size_t)n2);
    memset(b->cell2, 0, sizeof(int64_t) * (size_t)n3);
    memset(b->nodeIdTag, 0, sizeof(int64_t) * (size_t)n3);
    const int64_t *s = b->search;
    for (int64_t k = s[4]; k <= s[5]; ++k)
        for (int64_t j = s[2]; j <= s[3]; ++j)
            for (int64_t i = s[0]; i <= s[1]; ++i) tag_cell(b, i, j, k);
    for (int64_t k = s[4]; k <= s[5]; ++k)
        for (int64_t j = s[2]; j <= s[3]; ++j)
            for (int64_t i = s[0]; i <= s[1]; ++i)
                if (FLD(b, cell, i, j, k) != 2 && corner_sum(b, i, j, k) == 8) FLD(b, cell, i, j, k) = 1;
    count_tags(b);
}

void orc_selective_retagging(orc_block *b)
{
    if (b->blk_mv_tag != 0) return;
    int64_t I = b->ibCount;
    for (int64_t nn = 1; nn <= I; ++nn)
        FLD(b, cell, LIST(b->intercepted, I, nn, 1), LIST(b->intercepted, I, nn, 2),
            LIST(b->intercepted, I, nn, 3)) = 0;
    for (int64_t nn = 1; nn <= I; ++nn)
        for (int kp = -1; kp <= 1; ++kp)
            for (int jp = -1; jp <= 1; ++jp)
                for (int ip = -1; ip <= 1; ++ip)
                    tag_cell(b, LIST(b->intercepted, I, nn, 1) + ip, LIST(b->intercepted, I, nn, 2) + jp,
                             LIST(b->intercepted, I, nn, 3) + kp);
    for (int64_t nn = 1; nn <= I; ++nn) {
        int64_t i1 = LIST(b->intercepted, I, nn, 1), j1 = LIST(b->intercepted, I, nn, 2),
                k1 = LIST(b->intercepted, I, nn, 3);
        for (int64_t k = k1 - 1; k <= k1 + 1; ++k)
            for (int64_t j = j1 - 1; j <= j1 + 1; ++j)
                for (int64_t i = i1 - 1; i <= i1 + 1; ++i)
                    if (FLD(b, cell, i, j, k) != 2)
                        FLD(b, cell, i, j, k) = (corner_sum(b, i, j, k) == 8) ? 1 : 0;
    }
    count_tags(b);
}

static void split_red_black(orc_block *b)
{
    int64_t F = b->fluidCount, nr = 0, nb = 0;
    for (int64_t n = 1; n <= F; ++n) {
        int64_t s = LIST(b->fluid, F, n, 1) + LIST(b->fluid, F, n, 2) + LIST(b->fluid, F, n, 3);
        if (s % 2 == 1) nr++; else nb++;
    }
    b->redCount = nr;
    b->blackCount = nb;
    free(b->red); free(b->black);
    b->red = orc_xcalloc((size_t)(3 * nr + 1), sizeof(int64_t));
    b->black = orc_xcalloc((size_t)(3 * nb + 1), sizeof(int64_t));
    int64_t ir = 0, ib = 0;
    for (int64_t n = 1; n <= F; ++n) {
        int64_t i = LIST(b->fluid, F, n, 1), j = LIST(b->fluid, F, n, 2), k = LIST(b->fluid, F, n, 3);
        if ((i + j + k) % 2 == 1) {
            ++ir;
            LIST(b->red, nr, ir, 1) = i; LIST(b->red, nr, ir, 2) = j; LIST(b->red, nr, ir, 3) = k;
        } else {
            ++ib;
            LIST(b->black, nb, ib, 1) = i; LIST(b->black, nb, ib, 2) = j; LIST(b->black, nb, ib, 3) = k;
        }
    }
}

void orc_cell_count_solid(orc_block *b)
{
    int64_t F = b->fluidCount, S = b->solidCount, I = b->ibCount;
    free(b->fluid); free(b->solid); free(b->intercepted);
    b->fluid = orc_xcalloc((size_t)(3 * F + 1), sizeof(int64_t));
    b->solid = orc_xcalloc((size_t)(3 * S + 1), sizeof(int64_t));
    b->intercepted = orc_xcalloc((size_t)(3 * I + 1), sizeof(int64_t));
    int64_t a = 0, s = 0, c = 0;
    for (int64_t k = 2; k <= b->nz + 1; ++k)
        for (int64_t j = 2; j <= b->ny + 1; ++j)
            for (int64_t i = 2; i <= b->nx + 1; ++i) {
                int64_t v = FLD(b, cell, i, j, k);
                if (v == 0) { ++a; LIST(b->fluid, F, a, 1) = i; LIST(b->fluid, F, a, 2) = j; LIST(b->fluid, F, a, 3) = k; }
                else if (v == 1) { ++s; LIST(b->solid, S, s, 1) = i; LIST(b->solid, S, s, 2) = j; LIST(b->solid, S, s, 3) = k; }
                else if (v == 2) { ++c; LIST(b->intercepted, I, c, 1) = i; LIST(b->intercepted, I, c, 2) = j; LIST(b->intercepted, I, c, 3) = k; }
            }
    split_red_black(b);
}

void orc_cell_count_solid_coarse(orc_block *b)
{
    int64_t F = 0;
    for (int64_t k = 2; k <= b->nz + 1; ++k)
        for (int64_t j = 2; j <= b->ny + 1; ++j)
            for (int64_t i = 2; i <= b->nx + 1; ++i)
                if (FLD(b, cell, i, j, k) == 0) F++;
    b->fluidCount = F;
    free(b->fluid);
    b->fluid = orc_xcalloc((size_t)(3 * F + 1), sizeof(int64_t));
    int64_t a = 0;
    for (int64_t k = 2; k <= b->nz + 1; ++k)
        for (int64_t j = 2; j <= b->ny + 1; ++j)
            for (int64_t i = 2; i <= b->nx + 1; ++i)
                if (FLD(b, cell, i, j, k) == 0) {
                    ++a;
                    LIST(b->fluid, F, a, 1) = i; LIST(b->fluid, F, a, 2) = j; LIST(b->fluid, F, a, 3) = k;
                }
    split_red_black(b);
}

void orc_fine_block_cell(orc_case *c)
{
    orc_block *b1 = &c->blk[0];
    int64_t n2 = NX2(b1) * NY2(b1) * NZ2(b1);
    int64_t n3 = (b1->nx + 3) * (b1->ny + 3) * (b1->nz + 3);
    memset(b1->cell_n, 0, sizeof(int64_t) * (size_t)n2);
    memset(b1->cell, 0, sizeof(int64_t) * (size_t)n2);
    memset(b1->cell_pr, 0, sizeof(int64_t) * (size_t)n3);
    for (int64_t g = 1; g <= c->nintf; ++g) {
        orc_interface *it = &c->intf[g - 1];
        orc_block *a = &c->blk[it->a_blk - 1];
        orc_block *bb = &c->blk[it->b_blk - 1];
        double ci = (double)bb->cintp;
        for (int64_t k = 2; k <= a->nz + 1; ++k)
            for (int64_t j = 2; j <= a->ny + 1; ++j)
                for (int64_t i = 2; i <= a->nx + 1; ++i) {
                    double xp = V1(a->xp, i), yp = V1(a->yp, j), zp = V1(a->zp, k);
                    if (xp >= it->xs + ci * a->dx && xp <= it->xe - ci * a->dx &&
                        zp >= it->zs + ci * a->dz && zp <= it->ze - ci * a->dz &&
                        yp >= it->ys + ci * a->dy && yp <= it->ye - ci * a->dy) {
                        FLD(a, cell_n, i, j, k) = 1;
                        FLD(a, cell, i, j, k) = 1;
                    }
                    if (xp >= it->xs && xp <= it->xe && zp >= it->zs && zp <= it->ze &&
                        yp >= it->ys && yp <= it->ye)
                        TAG3(a, cell_pr, i, j, k) = 1;
                }
    }
}

void orc_compute_norm_distance(orc_block *b)
{
    int64_t I = b->ibCount;
    int64_t **nel[7] = { &b->nelp, &b->nelu1, &b->nelu2, &b->nelv1, &b->nelv2, &b->nelw1, &b->nelw2 };
    double **nd[7] = { &b->pNormDis, &b->u1NormDis, &b->u2NormDis, &b->v1NormDis, &b->v2NormDis,
                       &b->w1NormDis, &b->w2NormDis };
    for (int q = 0; q < 7; ++q) {
        free(*nel[q]); free(*nd[q]);
        *nel[q] = orc_xcalloc((size_t)I + 1, sizeof(int64_t));
        *nd[q] = orc_xcalloc((size_t)I + 1, sizeof(double));
    }
    for (int64_t k = 1; k <= I; ++k) {
        int64_t ii = LIST(b->intercepted, I, k, 1), jj = LIST(b->intercepted, I, k, 2),
                kk = LIST(b->intercepted, I, k, 3);
        double n1x = V1(b->xp, ii), n2x = V1(b->x1, ii), n3x = V1(b->x1, ii + 1);
        double n1y = V1(b->yp, jj), n2y = V1(b->y1, jj), n3y = V1(b->y1, jj + 1);
        double n1z = V1(b->zp, kk), n2z = V1(b->z1, kk), n3z = V1(b->z1, kk + 1);
        double md[7];
        int64_t mi[7] = { 0, 0, 0, 0, 0, 0, 0 };
        for (int q = 0; q < 7; ++q) md[q] = 1e14;
        for (int64_t m = 1; m <= b->ibElems; ++m) {
            double cx = V1(b->xcent, m), cy = V1(b->ycent, m), cz = V1(b->zcent, m);
            double d[7];
            d[0] = sq(n1x - cx) + sq(n1y - cy) + sq(n1z - cz);
            d[1] = sq(n2x - cx) + sq(n1y - cy) + sq(n1z - cz);
            d[2] = sq(n3x - cx) + sq(n1y - cy) + sq(n1z - cz);
            d[3] = sq(n1x - cx) + sq(n2y - cy) + sq(n1z - cz);
            d[4] = sq(n1x - cx) + sq(n3y - cy) + sq(n1z - cz);
            d[5] = sq(n1x - cx) + sq(n1y - cy) + sq(n2z - cz);
            d[6] = sq(n1x - cx) + sq(n1y - cy) + sq(n3z - cz);
            for (int q = 0; q < 7; ++q)
                if (d[q] < md[q]) { md[q] = d[q]; mi[q] = m; }
        }
        /* sample points of the seven distances: p, u1, u2, v1, v2, w1, w2 */
        double px[7] = { n1x, n2x, n3x, n1x, n1x, n1x, n1x };
        double py[7] = { n1y, n1y, n1y, n2y, n3y, n1y, n1y };
        double pz[7] = { n1z, n1z, n1z, n1z, n1z, n2z, n3z };
        for (int q = 0; q < 7; ++q) {
            int64_t e = mi[q];
            V1(*nel[q], k) = e;
            V1(*nd[q], k) = (px[q] - V1(b->xcent, e)) * V1(b->cosA, e) +
                            (py[q] - V1(b->ycent, e)) * V1(b->cosB, e) +
                            (pz[q] - V1(b->zcent, e)) * V1(b->cosG, e);
        }
    }
}

void orc_find_ts_cells(orc_block *b)
{
    for (int64_t k = 2; k <= b->nz + 1; ++k)
        for (int64_t j = 2; j <= b->ny + 1; ++j)
            for (int64_t i = 2; i <= b->nx + 1; ++i) TAG3(b, cell2, i, j, k) = 0;
    int64_t I = b->ibCount, t = 0;
    for (int64_t n = 1; n <= I; ++n) {
        int64_t sid = V1(b->surfId, V1(b->nelp, n));
        if (sid == 51 || sid == 52) {
            TAG3(b, cell2, LIST(b->intercepted, I, n, 1), LIST(b->intercepted, I, n, 2),
                 LIST(b->intercepted, I, n, 3)) = 2;
            t++;
        }
    }
    b->tsCount = t;
    free(b->ts); free(b->index_ts);
    b->ts = orc_xcalloc((size_t)(3 * t + 1), sizeof(int64_t));
    b->index_ts = orc_xcalloc((size_t)t + 1, sizeof(int64_t));
    int64_t a = 0;
    for (int64_t n = 1; n <= I; ++n) {
        int64_t i = LIST(b->intercepted, I, n, 1), j = LIST(b->intercepted, I, n, 2),
                k = LIST(b->intercepted, I, n, 3);
        if (TAG3(b, cell2, i, j, k) == 2) {
            ++a;
            LIST(b->ts, t, a, 1) = i; LIST(b->ts, t, a, 2) = j; LIST(b->ts, t, a, 3) = k;
            V1(b->index_ts, a) = n;
        }
    }
    for (int q = 0; q < ORC_NGHOST; ++q) {
        free(b->ghost[q]);
        b->ghost[q] = orc_xcalloc((size_t)t + 1, sizeof(double));
    }
}
