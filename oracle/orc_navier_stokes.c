/*
 * Test infrastructure: FP64 restatement of the reference momentum step,
 * reference src/NavierStokes.f90 (non_uni_coeff :13-250, nsMomentum2order
 * :252-912, compute_f/s/ak :915-996, check_adjacent_cell :1002-1022) and of
 * coefficientMatrix (src/CoefficientMatrix.f90:10-71).
 * Expression order follows the Fortran source (left-to-right sums, no FMA).
 */
#include "orc_internal.h"

/* ---------------- CoefficientMatrix.f90:10-71 ---------------- */
static void coef_axis(const double *cc, int64_t n, double *A)
{
    for (int64_t j = 2; j <= n + 1; ++j) {
        double r1 = V1(cc, j) - V1(cc, j - 1);
        double r2 = V1(cc, j + 1) - V1(cc, j);
        double rs = r1 + r2;
        AC(A, n, j - 1, 1) = 2.0 / (r1 * rs);
        AC(A, n, j - 1, 2) = -2.0 / (r1 * r2);
        AC(A, n, j - 1, 3) = 2.0 / (r2 * rs);
    }
}

void orc_coefficient_matrix(orc_block *b, int is_coarse)
{
    coef_axis(b->yp, b->ny, b->Acy);
    coef_axis(b->xp, b->nx, b->Acx);
    coef_axis(b->zp, b->nz, b->Acz);
    if (is_coarse) {
        int64_t nx = b->nx, ny = b->ny, nz = b->nz;
        AC(b->Acx, nx, 1, 2) = AC(b->Acx, nx, 1, 2) - AC(b->Acx, nx, 1, 1);
        AC(b->Acx, nx, 1, 1) = 0.0;
        AC(b->Acx, nx, nx, 2) = AC(b->Acx, nx, nx, 2) - AC(b->Acx, nx, nx, 3);
        AC(b->Acx, nx, nx, 3) = 0.0;
        AC(b->Acy, ny, 1, 2) = AC(b->Acy, ny, 1, 2) + AC(b->Acy, ny, 1, 1);
        AC(b->Acy, ny, 1, 1) = 0.0;
        AC(b->Acy, ny, ny, 2) = AC(b->Acy, ny, ny, 2) + AC(b->Acy, ny, ny, 3);
        AC(b->Acy, ny, ny, 3) = 0.0;
        AC(b->Acz, nz, 1, 2) = AC(b->Acz, nz, 1, 2) + AC(b->Acz, nz, 1, 1);
        AC(b->Acz, nz, 1, 1) = 0.0;
        AC(b->Acz, nz, nz, 2) = AC(b->Acz, nz, nz, 2) + AC(b->Acz, nz, nz, 3);
        AC(b->Acz, nz, nz, 3) = 0.0;
    }
}

/* ---------------- NavierStokes.f90:915-996 ---------------- */
static double ipow(double x, int n)
{
    /* small constant integer powers as the Fortran compiler expands them */
    switch (n) {
    case 2: return x * x;
    case 3: return x * x * x;
    case 4: { double x2 = x * x; return x2 * x2; }
    default: return pow(x, (double)n);
    }
}

static void compute_f(const double th[3], double f[9])
{
    f[1] = th[2];
    f[2] = 2.0 * th[2] + ipow(th[2], 2);
    f[3] = 3.0 * th[2] + 3.0 * ipow(th[2], 2) + ipow(th[2], 3);
    f[4] = 4.0 * th[2] + 6.0 * ipow(th[2], 2) + 4.0 * ipow(th[2], 3) + ipow(th[2], 4);
    f[5] = 1.0 / (th[0] * th[1]);
    f[6] = (2.0 * th[0] + 1.0) / (ipow(th[0], 2) * ipow(th[1], 2));
    f[7] = (3.0 * ipow(th[0], 2) + 3.0 * th[0] + 1.0) / (ipow(th[0], 3) * ipow(th[1], 3));
    f[8] = (4.0 * ipow(th[0], 3) + 6.0 * ipow(th[0], 2) + 4.0 * th[0] + 1.0) /
           (ipow(th[0], 4) * ipow(th[1], 4));
}

/* s(5:10, 1:7) stored as s[row][col] with row 5..10, col 1..7 */
static void compute_s(double t2, const double f[9], double s[11][8])
{
    s[5][1] = -1.0 / ipow(t2, 4);
    s[5][2] = f[4];
    s[5][3] = f[4];
    s[5][4] = s[5][1];
    s[5][5] = -(f[1] / ipow(t2, 4) + f[4] / t2);
    s[5][6] = (f[4] / ipow(t2, 2) - f[2] / ipow(t2, 4));
    s[5][7] = -(f[3] / ipow(t2, 4) + f[4] / ipow(t2, 3));

    s[6][1] = -1.0;
    s[6][2] = f[8];
    s[6][3] = f[8];
    s[6][4] = -1.0;
    s[6][5] = (f[5] + f[8]);
    s[6][6] = (f[8] - f[6]);
    s[6][7] = (f[7] + f[8]);

    s[7][1] = -1.0;
    s[7][2] = (1.0 + f[4]);
    s[7][3] = f[4];
    s[7][4] = (f[4] - f[1]);
    s[7][5] = (f[4] - f[2]);
    s[7][6] = (f[4] - f[3]);

    s[8][1] = -1.0 / pow(t2, 4.0);
    s[8][2] = (f[8] + (1.0 / pow(t2, 4.0)));
    s[8][3] = f[8];
    s[8][4] = (f[5] / pow(t2, 4.0) - f[8] / t2);
    s[8][5] = (f[8] / pow(t2, 2.0) - f[6] / pow(t2, 4.0));
    s[8][6] = (f[7] / pow(t2, 4.0) - f[8] / pow(t2, 3.0));

    s[9][1] = -s[6][6] * s[5][1];
    s[9][2] = s[5][6] * s[6][1];
    s[9][3] = -(s[5][4] * s[6][6] + s[5][6] * s[6][2]);
    s[9][4] = (s[5][6] * s[6][4] + s[6][6] * s[5][2]);
    s[9][5] = (s[5][6] * s[6][3] - s[6][6] * s[5][3]);
    s[9][6] = (s[5][6] * s[6][5] - s[6][6] * s[5][5]);
    s[9][7] = (s[5][6] * s[6][7] - s[6][6] * s[5][7]);

    s[10][1] = -s[8][5] * s[7][1];
    s[10][2] = -s[8][5] * s[7][2];
    s[10][3] = s[7][5] * s[8][1];
    s[10][4] = s[7][5] * s[8][2];
    s[10][5] = (s[7][5] * s[8][3] - s[8][5] * s[7][3]);
    s[10][6] = (s[7][5] * s[8][4] - s[8][5] * s[7][4]);
    s[10][7] = (s[7][5] * s[8][6] - s[8][5] * s[7][6]);
}

static void compute_ak(const double th[3], double ak[8])
{
    ak[1] = (1.0 + 2.0 * th[0]) * (th[2] + ipow(th[2], 2)) * th[1];
    ak[2] = (1.0 + th[0]) * (2.0 * th[2] + ipow(th[2], 2));
    ak[3] = (1.0 + th[0]);
    ak[4] = (1.0 + th[0]) * ipow(1.0 + th[2], 2);
    ak[5] = ipow(1.0 + th[0], 2) * (th[2] + ipow(th[2], 2)) * th[1];
    ak[6] = ipow(th[0], 2) * th[1] * (th[2] + ipow(th[2], 2));
    ak[7] = (1.0 + th[0]) * (th[2] + ipow(th[2], 2)) * th[1];
}

static void fill_pair(const double th[3], double *ca, double *ck, int64_t i)
{
    double f[9], s[11][8], ak[8];
    compute_f(th, f);
    compute_s(th[1], f, s);
    TAB(ca, 1, i) = s[9][7] * s[10][1] - s[10][7] * s[9][1];
    TAB(ca, 2, i) = s[9][7] * s[10][2] + s[10][7] * s[9][3];
    TAB(ca, 3, i) = s[9][5] * s[10][7] - s[10][5] * s[9][7];
    TAB(ca, 4, i) = s[9][4] * s[10][7] + s[10][4] * s[9][7];
    TAB(ca, 5, i) = s[9][7] * s[10][3] - s[10][7] * s[9][2];
    TAB(ca, 6, i) = s[9][7] * s[10][6] - s[10][7] * s[9][6];
    compute_ak(th, ak);
    TAB(ck, 1, i) = ak[3];
    TAB(ck, 2, i) = -ak[4];
    TAB(ck, 3, i) = ak[1] + ak[2];
    TAB(ck, 4, i) = -ak[5];
    TAB(ck, 5, i) = ak[6];
    TAB(ck, 6, i) = ak[7];
}

static void uni_axis_same(const double *d, int64_t n, double *ca, double *ck)
{
    for (int64_t i = 2; i <= n; ++i) {
        double th[3] = { V1(d, i) / V1(d, i - 1), V1(d, i + 1) / V1(d, i), V1(d, i + 2) / V1(d, i + 1) };
        fill_pair(th, ca, ck, i);
    }
}

static void uni_axis_cross(const double *d, int64_t n, double *ca, double *ck)
{
    for (int64_t i = 2; i <= n; ++i) {
        double t1;
        if (i == 2) t1 = V1(d, i - 1);
        else t1 = 0.5 * (V1(d, i - 1) + V1(d, i - 2));
        double t2 = 0.5 * (V1(d, i) + V1(d, i - 1));
        double t3 = 0.5 * (V1(d, i) + V1(d, i + 1));
        double t4 = 0.5 * (V1(d, i + 1) + V1(d, i + 2));
        double th[3] = { t2 / t1, t3 / t2, t4 / t3 };
        fill_pair(th, ca, ck, i);
    }
}

void orc_non_uni_coeff(orc_block *b)
{
    int64_t ns[T_COUNT] = { b->nx, b->ny, b->nz, b->nx, b->nx, b->ny, b->ny, b->nz, b->nz };
    for (int t = 0; t < T_COUNT; ++t) {
        free(b->ca[t]);
        free(b->ck[t]);
        b->ca[t] = orc_xcalloc((size_t)(6 * (ns[t] + 2)), sizeof(double));
        b->ck[t] = orc_xcalloc((size_t)(6 * (ns[t] + 2)), sizeof(double));
    }
    uni_axis_same(b->deltax, b->nx, b->ca[T_UU], b->ck[T_UU]);
    uni_axis_same(b->deltay, b->ny, b->ca[T_VV], b->ck[T_VV]);
    uni_axis_same(b->deltaz, b->nz, b->ca[T_WW], b->ck[T_WW]);
    uni_axis_cross(b->deltax, b->nx, b->ca[T_UV], b->ck[T_UV]);
    uni_axis_cross(b->deltay, b->ny, b->ca[T_VU], b->ck[T_VU]);
    uni_axis_cross(b->deltaz, b->nz, b->ca[T_WU], b->ck[T_WU]);
    /* ca_uw = ca_uv, ca_vw = ca_vu, ca_wv = ca_wu (NavierStokes.f90:134-155, 180-201, 226-247) */
    for (int64_t q = 0; q < 6 * (b->nx + 2); ++q) { b->ca[T_UW][q] = b->ca[T_UV][q]; b->ck[T_UW][q] = b->ck[T_UV][q]; }
    for (int64_t q = 0; q < 6 * (b->ny + 2); ++q) { b->ca[T_VW][q] = b->ca[T_VU][q]; b->ck[T_VW][q] = b->ck[T_VU][q]; }
    for (int64_t q = 0; q < 6 * (b->nz + 2); ++q) { b->ca[T_WV][q] = b->ca[T_WU][q]; b->ck[T_WV][q] = b->ck[T_WU][q]; }
    /* entries outside 2..n stay 0: Fortran leaves them unset; never read */
}

/* ---------------- NavierStokes.f90:1002-1022 ---------------- */
static int adj_ok(const orc_block *b, const int64_t *cell, int tall, int64_t i, int64_t j,
                  int64_t k, int step)
{
#define CV(ii, jj, kk) (tall ? cell[IXT(b, ii, jj, kk)] : cell[IX(b, ii, jj, kk)])
    if (CV(i - step, j, k) == 2) return 0;
    if (CV(i + step, j, k) == 2) return 0;
    if (CV(i, j - step, k) == 2) return 0;
    if (CV(i, j + step, k) == 2) return 0;
    if (CV(i, j, k - step) == 2) return 0;
    if (CV(i, j, k + step) == 2) return 0;
    return 1;
#undef CV
}

/* 5-point third-order upwind term (NavierStokes.f90:619-647 pattern):
 * vel*(ca1 f(+2) + ca2 f(+1) + ca3 f0 + ca4 f(-1) + ca5 f(-2))/(ca6*hA)
 *  + |vel|*(ck1 f(+2) + ... + ck5 f(-2))/(2 ck6 hB) */
static double upw3(double vel, const double *ca, const double *ck, int64_t t, double fp2,
                   double fp1, double f0, double fm1, double fm2, double hA, double hB)
{
    return vel * (TAB(ca, 1, t) * fp2 + TAB(ca, 2, t) * fp1 + TAB(ca, 3, t) * f0 +
                  TAB(ca, 4, t) * fm1 + TAB(ca, 5, t) * fm2) /
               (TAB(ca, 6, t) * hA) +
           fabs(vel) * (TAB(ck, 1, t) * fp2 + TAB(ck, 2, t) * fp1 + TAB(ck, 3, t) * f0 +
                        TAB(ck, 4, t) * fm1 + TAB(ck, 5, t) * fm2) /
               (2.0 * TAB(ck, 6, t) * hB);
}

static double dotn(const orc_block *b, int64_t idx, double x, double y, double z)
{
    return (x - V1(b->xcent, idx)) * V1(b->cosA, idx) + (y - V1(b->ycent, idx)) * V1(b->cosB, idx) +
           (z - V1(b->zcent, idx)) * V1(b->cosG, idx);
}

void orc_ns_momentum(orc_block *b, double dt, double re)
{
    const double alpha = 1.0;
    int64_t nx = b->nx, ny = b->ny, nz = b->nz;
    double *u = b->u, *v = b->v, *w = b->w, *p = b->p;
#define U(i, j, k) u[IX(b, i, j, k)]
#define Vv(i, j, k) v[IX(b, i, j, k)]
#define W(i, j, k) w[IX(b, i, j, k)]
#define P(i, j, k) p[IX(b, i, j, k)]
#define C2(i, j, k) b->cell2[IXT(b, i, j, k)]
    for (int64_t n = 1; n <= b->fluidCount; ++n) {
        int64_t i = LIST(b->fluid, b->fluidCount, n, 1);
        int64_t j = LIST(b->fluid, b->fluidCount, n, 2);
        int64_t k = LIST(b->fluid, b->fluidCount, n, 3);
        int64_t iip1 = 0, iim1 = 0, ijp1 = 0, ijm1 = 0, ikp1 = 0, ikm1 = 0;
        if (C2(i + 1, j, k) == 2 || C2(i - 1, j, k) == 2 || C2(i, j + 1, k) == 2 ||
            C2(i, j - 1, k) == 2 || C2(i, j, k + 1) == 2 || C2(i, j, k - 1) == 2) {
            for (int64_t nn = 1; nn <= b->tsCount; ++nn) {
                int64_t bt = V1(b->index_ts, nn);
                double tu2 = dotn(b, V1(b->nelu2, bt), V1(b->x1, i + 1), V1(b->yp, j), V1(b->zp, k));
                double tu1 = dotn(b, V1(b->nelu1, bt), V1(b->x1, i), V1(b->yp, j), V1(b->zp, k));
                double tv2 = dotn(b, V1(b->nelv2, bt), V1(b->xp, i), V1(b->y1, j + 1), V1(b->zp, k));
                double tv1 = dotn(b, V1(b->nelv1, bt), V1(b->xp, i), V1(b->y1, j), V1(b->zp, k));
                double tw2 = dotn(b, V1(b->nelw2, bt), V1(b->xp, i), V1(b->yp, j), V1(b->z1, k + 1));
                double tw1 = dotn(b, V1(b->nelw1, bt), V1(b->xp, i), V1(b->yp, j), V1(b->z1, k));
                double tp = dotn(b, V1(b->nelp, bt), V1(b->xp, i), V1(b->yp, j), V1(b->zp, k));
                double *g = NULL;
#define GH(q) V1(b->ghost[q], nn)
                (void)g;
                if (C2(i + 1, j, k) == 2) {
                    iip1 = nn;
                    if (tu2 < 0) U(i + 1, j, k) = GH(G_U2);
                    if (tv2 < 0) Vv(i + 1, j, k) = GH(G_V2);
                    if (tv1 < 0) Vv(i + 1, j - 1, k) = GH(G_V1);
                    if (tw2 < 0) W(i + 1, j, k) = GH(G_W2);
                    if (tw1 < 0) W(i + 1, j, k - 1) = GH(G_W1);
                    if (tp < 0) P(i + 1, j, k) = GH(G_P);
                }
                if (C2(i - 1, j, k) == 2) {
                    iim1 = nn;
                    if (tu1 < 0) U(i - 2, j, k) = GH(G_U1);
                    if (tv2 < 0) Vv(i - 1, j, k) = GH(G_V2);
                    if (tv1 < 0) Vv(i - 1, j - 1, k) = GH(G_V1);
                    if (tw2 < 0) W(i - 1, j, k) = GH(G_W2);
                    if (tw1 < 0) W(i - 1, j, k - 1) = GH(G_W1);
                    if (tp < 0) P(i - 1, j, k) = GH(G_P);
                }
                if (C2(i, j + 1, k) == 2) {
                    ijp1 = nn;
                    if (tu2 < 0) U(i, j + 1, k) = GH(G_U2);
                    if (tu1 < 0) U(i - 1, j + 1, k) = GH(G_U1);
                    if (tv2 < 0) Vv(i, j + 1, k) = GH(G_V2);
                    if (tw2 < 0) W(i, j + 1, k) = GH(G_W2);
                    if (tw1 < 0) W(i, j + 1, k - 1) = GH(G_W1);
                    if (tp < 0) P(i, j + 1, k) = GH(G_P);
                }
                if (C2(i, j - 1, k) == 2) {
                    ijm1 = nn;
                    if (tu2 < 0) U(i, j - 1, k) = GH(G_U2);
                    if (tu1 < 0) U(i - 1, j - 1, k) = GH(G_U1);
                    if (tv1 < 0) Vv(i, j - 2, k) = GH(G_V1);
                    if (tw2 < 0) W(i, j - 1, k) = GH(G_W2);
                    if (tw1 < 0) W(i, j - 1, k - 1) = GH(G_W1);
                    if (tp < 0) P(i, j - 1, k) = GH(G_P);
                }
                if (C2(i, j, k + 1) == 2) {
                    ikp1 = nn;
                    if (tu2 < 0) U(i, j, k + 1) = GH(G_U2);
                    if (tu1 < 0) U(i - 1, j, k + 1) = GH(G_U1);
                    if (tv2 < 0) Vv(i, j, k + 1) = GH(G_V2);
                    if (tv1 < 0) Vv(i, j - 1, k + 1) = GH(G_V1);
                    if (tw2 < 0) W(i, j, k + 1) = GH(G_W2);
                    if (tp < 0) P(i, j, k + 1) = GH(G_P);
                }
                if (C2(i, j, k - 1) == 2) {
                    ikm1 = nn;
                    if (tu2 < 0) U(i, j, k - 1) = GH(G_U2);
                    if (tu1 < 0) U(i - 1, j, k - 1) = GH(G_U1);
                    if (tv2 < 0) Vv(i, j, k - 1) = GH(G_V2);
                    if (tv1 < 0) Vv(i, j - 1, k - 1) = GH(G_V1);
                    if (tw1 < 0) W(i, j, k - 2) = GH(G_W1);
                    if (tp < 0) P(i, j, k - 1) = GH(G_P);
                }
#undef GH
            }
        }
        double dxr = V1(b->deltax, i + 1), dx = V1(b->deltax, i), dxl = V1(b->deltax, i - 1);
        double dye = V1(b->deltay, j + 1), dy = V1(b->deltay, j), dyw = V1(b->deltay, j - 1);
        double dzt = V1(b->deltaz, k + 1), dz = V1(b->deltaz, k), dzb = V1(b->deltaz, k - 1);

        double u1a = U(i - 1, j, k) + U(i, j, k);
        double u22 = U(i - 1, j, k) - U(i, j, k);
        double u3 = U(i, j, k) + U(i + 1, j, k);
        double u4 = U(i, j, k) - U(i + 1, j, k);
        double u5 = U(i, j - 1, k) + U(i, j, k);
        double u6 = U(i, j - 1, k) - U(i, j, k);
        double u7 = U(i, j, k) + U(i, j + 1, k);
        double u8 = U(i, j, k) - U(i, j + 1, k);
        double u9 = U(i, j, k - 1) + U(i, j, k);
        double u10 = U(i, j, k - 1) - U(i, j, k);
        double u11 = U(i, j, k) + U(i, j, k + 1);
        double u12 = U(i, j, k) - U(i, j, k + 1);
        double u13 = U(i - 1, j, k) + U(i - 1, j + 1, k);
        double u14 = u7;
        double u15 = U(i - 1, j, k) + U(i - 1, j, k + 1);
        double u16 = u11;

        double v1a = Vv(i, j - 1, k) + Vv(i + 1, j - 1, k);
        double v22 = Vv(i, j, k) + Vv(i + 1, j, k);
        double v3 = Vv(i - 1, j, k) + Vv(i, j, k);
        double v4 = Vv(i - 1, j, k) - Vv(i, j, k);
        double v5 = Vv(i, j, k) + Vv(i + 1, j, k);
        double v6 = Vv(i, j, k) - Vv(i + 1, j, k);
        double v7 = Vv(i, j - 1, k) + Vv(i, j, k);
        double v8 = Vv(i, j - 1, k) - Vv(i, j, k);
        double v9 = Vv(i, j, k) + Vv(i, j + 1, k);
        double v10 = Vv(i, j, k) - Vv(i, j + 1, k);
        double v11 = Vv(i, j, k - 1) + Vv(i, j, k);
        double v12 = Vv(i, j, k - 1) - Vv(i, j, k);
        double v13 = Vv(i, j, k) + Vv(i, j, k + 1);
        double v14 = Vv(i, j, k) - Vv(i, j, k + 1);
        double v15 = Vv(i, j - 1, k) + Vv(i, j - 1, k + 1);
        double v16 = v13;

        double w1a = W(i, j, k - 1) + W(i + 1, j, k - 1);
        double w22 = W(i, j, k) + W(i + 1, j, k);
        double w3 = W(i, j, k - 1) + W(i, j + 1, k - 1);
        double w4 = W(i, j, k) + W(i, j + 1, k);
        double w5 = W(i - 1, j, k) + W(i, j, k);
        double w6 = W(i - 1, j, k) - W(i, j, k);
        double w7 = W(i, j, k) + W(i + 1, j, k);
        double w8 = W(i, j, k) - W(i + 1, j, k);
        double w9 = W(i, j - 1, k) + W(i, j, k);
        double w10 = W(i, j - 1, k) - W(i, j, k);
        double w11 = W(i, j, k) + W(i, j + 1, k);
        double w12 = W(i, j, k) - W(i, j + 1, k);
        double w13 = W(i, j, k - 1) + W(i, j, k);
        double w14 = W(i, j, k - 1) - W(i, j, k);
        double w15 = W(i, j, k) + W(i, j, k + 1);
        double w16 = W(i, j, k) - W(i, j, k + 1);

        double dpdx = (P(i, j, k) - P(i + 1, j, k)) / (0.5 * (dxr + dx));
        double dpdy = (P(i, j, k) - P(i, j + 1, k)) / (0.5 * (dye + dy));
        double dpdz = (P(i, j, k) - P(i, j, k + 1)) / (0.5 * (dzt + dz));

        double dx2xr = dx + dxr, dx2xl = dx + dxl;
        double dy2ye = dy + dye, dy2yw = dy + dyw;
        double dz2zt = dz + dzt, dz2zb = dz + dzb;

        /* ---- u momentum ---- */
        double vu_e = Vv(i + 1, j, k) + (dxr / (dx2xr)) * (Vv(i, j, k) - Vv(i + 1, j, k));
        double vu_w = Vv(i + 1, j - 1, k) + (dxr / (dx2xr)) * (Vv(i, j - 1, k) - Vv(i + 1, j - 1, k));
        double v_in_um = 0.5 * (vu_e + vu_w);
        double wu_n = W(i + 1, j, k) + (dxr / (dx2xr)) * (W(i, j, k) - W(i + 1, j, k));
        double wu_s = W(i + 1, j, k - 1) + (dxr / (dx2xr)) * (W(i, j, k - 1) - W(i + 1, j, k - 1));
        double w_in_um = 0.5 * (wu_n + wu_s);
        int third = 0;
        if (i > 2 && i < nx && j > 2 && j < ny + 1 && k > 2 && k < nz + 1)
            if (adj_ok(b, b->cell, 0, i, j, k, 1) && adj_ok(b, b->cell2, 1, i, j, k, 2)) third = 1;
        double duudx, dvudy, dwudz;
        if (third) {
            double ddy = 0.5 * (V1(b->deltay, j) + V1(b->deltay, j - 1));
            double ddye = 0.5 * (V1(b->deltay, j) + V1(b->deltay, j + 1));
            double ddz = 0.5 * (V1(b->deltaz, k) + V1(b->deltaz, k - 1));
            double ddzr = 0.5 * (V1(b->deltaz, k) + V1(b->deltaz, k + 1));
            duudx = upw3(U(i, j, k), b->ca[T_UU], b->ck[T_UU], i, U(i + 2, j, k), U(i + 1, j, k),
                         U(i, j, k), U(i - 1, j, k), U(i - 2, j, k), V1(b->deltax, i + 1), V1(b->deltax, i));
            dvudy = upw3(v_in_um, b->ca[T_VU], b->ck[T_VU], j, U(i, j + 2, k), U(i, j + 1, k),
                         U(i, j, k), U(i, j - 1, k), U(i, j - 2, k), ddye, ddy);
            dwudz = upw3(w_in_um, b->ca[T_WU], b->ck[T_WU], k, U(i, j, k + 2), U(i, j, k + 1),
                         U(i, j, k), U(i, j, k - 1), U(i, j, k - 2), ddzr, ddz);
        } else {
            duudx = -0.25 * (u1a * u1a + alpha * fabs(u1a) * u22 - u3 * u3 - alpha * fabs(u3) * u4) / V1(b->deltax, i);
            dvudy = -0.25 * (v1a * u5 + alpha * fabs(v1a) * u6 - v22 * u7 - alpha * fabs(v22) * u8) / V1(b->deltay, j);
            dwudz = -0.25 * (w1a * u9 + alpha * fabs(w1a) * u10 - w22 * u11 - alpha * fabs(w22) * u12) / V1(b->deltaz, k);
        }
        double d2udx2 = (2.0 / dx2xr) * ((-u4 / dxr) + (u22 / dx));
        double d2udy2 = (2.0 / (0.5 * (dy2ye + dy2yw))) * ((-u8 / (0.5 * dy2ye)) + (u6 / (0.5 * dy2yw)));
        double d2udz2 = (2.0 / (0.5 * (dz2zt + dz2zb))) * ((-u12 / (0.5 * dz2zt)) + (u10 / (0.5 * dz2zb)));
        double xtt2 = (d2udx2 + d2udy2 + d2udz2) / re;
        double residu = (-duudx - dvudy - dwudz + xtt2);
        FLD(b, ut, i, j, k) = U(i, j, k) + dt * (residu + dpdx);

        /* ---- v momentum ---- */
        double uv_e = U(i, j + 1, k) + (dye / (dy2ye)) * (U(i, j, k) - U(i, j + 1, k));
        double uv_w = U(i - 1, j + 1, k) + (dye / (dy2ye)) * (U(i - 1, j, k) - U(i - 1, j + 1, k));
        double u_in_vm = 0.5 * (uv_e + uv_w);
        double wv_n = W(i, j + 1, k) + (dye / (dy2ye)) * (W(i, j, k) - W(i, j + 1, k));
        double wv_s = W(i, j + 1, k - 1) + (dye / (dy2ye)) * (W(i, j, k - 1) - W(i, j + 1, k - 1));
        double w_in_vm = 0.5 * (wv_n + wv_s);
        third = 0;
        if (i > 2 && i < nx + 1 && j > 2 && j < ny && k > 2 && k < nz + 1)
            if (adj_ok(b, b->cell, 0, i, j, k, 1) && adj_ok(b, b->cell2, 1, i, j, k, 2)) third = 1;
        double duvdx, dvvdy, dwvdz;
        if (third) {
            double ddx = 0.5 * (V1(b->deltax, i) + V1(b->deltax, i - 1));
            double ddxr = 0.5 * (V1(b->deltax, i) + V1(b->deltax, i + 1));
            double ddz = 0.5 * (V1(b->deltaz, k) + V1(b->deltaz, k - 1));
            double ddzr = 0.5 * (V1(b->deltaz, k) + V1(b->deltaz, k + 1));
            duvdx = upw3(u_in_vm, b->ca[T_UW], b->ck[T_UW], i, Vv(i + 2, j, k), Vv(i + 1, j, k),
                         Vv(i, j, k), Vv(i - 1, j, k), Vv(i - 2, j, k), ddxr, ddx);
            dvvdy = upw3(Vv(i, j, k), b->ca[T_VV], b->ck[T_VV], j, Vv(i, j + 2, k), Vv(i, j + 1, k),
                         Vv(i, j, k), Vv(i, j - 1, k), Vv(i, j - 2, k), V1(b->deltay, j + 1), V1(b->deltay, j));
            dwvdz = upw3(w_in_vm, b->ca[T_WU], b->ck[T_WU], k, Vv(i, j, k + 2), Vv(i, j, k + 1),
                         Vv(i, j, k), Vv(i, j, k - 1), Vv(i, j, k - 2), ddzr, ddz);
        } else {
            duvdx = -0.25 * (u13 * v3 + alpha * fabs(u13) * v4 - u14 * v5 - alpha * fabs(u14) * v6) / V1(b->deltax, i);
            dvvdy = -0.25 * (v7 * v7 + alpha * fabs(v7) * v8 - v9 * v9 - alpha * fabs(v9) * v10) / V1(b->deltay, j);
            dwvdz = -0.25 * (w3 * v11 + alpha * fabs(w3) * v12 - w4 * v13 - alpha * fabs(w4) * v14) / V1(b->deltaz, k);
        }
        double d2vdx2 = (2.0 / (0.5 * (dx2xr + dx2xl))) * ((-v6 / (0.5 * dx2xr)) + (v4 / (0.5 * dx2xl)));
        double d2vdy2 = (2.0 / dy2ye) * ((-v10 / dye) + (v8 / dy));
        double d2vdz2 = (2.0 / (0.5 * (dz2zt + dz2zb))) * ((-v14 / (0.5 * dz2zt)) + (v12 / (0.5 * dz2zb)));
        double ytt2 = (d2vdx2 + d2vdy2 + d2vdz2) / re;
        double residv = (-duvdx - dvvdy - dwvdz + ytt2);
        FLD(b, vt, i, j, k) = Vv(i, j, k) + dt * (residv + dpdy);

        /* ---- w momentum ---- */
        double uw_e = U(i, j, k + 1) + (dzt / (dz2zt)) * (U(i, j, k) - U(i, j, k + 1));
        double uw_w = U(i - 1, j, k + 1) + (dzt / (dz2zt)) * (U(i - 1, j, k) - U(i - 1, j, k + 1));
        double u_in_wm = 0.5 * (uw_e + uw_w);
        double vw_n = Vv(i, j, k + 1) + (dzt / (dz2zt)) * (Vv(i, j, k) - Vv(i, j, k + 1));
        double vw_s = Vv(i, j - 1, k + 1) + (dzt / (dz2zt)) * (Vv(i, j - 1, k) - Vv(i, j - 1, k + 1));
        double v_in_wm = 0.5 * (vw_n + vw_s);
        third = 0;
        if (i > 2 && i < nx + 1 && j > 2 && j < ny + 1 && k > 2 && k < nz)
            if (adj_ok(b, b->cell, 0, i, j, k, 1) && adj_ok(b, b->cell2, 1, i, j, k, 2)) third = 1;
        double duwdx, dvwdy, dwwdz;
        if (third) {
            double ddx = 0.5 * (V1(b->deltax, i) + V1(b->deltax, i - 1));
            double ddxr = 0.5 * (V1(b->deltax, i) + V1(b->deltax, i + 1));
            double ddy = 0.5 * (V1(b->deltay, j) + V1(b->deltay, j - 1));
            double ddye = 0.5 * (V1(b->deltay, j) + V1(b->deltay, j + 1));
            duwdx = upw3(u_in_wm, b->ca[T_UW], b->ck[T_UW], i, W(i + 2, j, k), W(i + 1, j, k),
                         W(i, j, k), W(i - 1, j, k), W(i - 2, j, k), ddxr, ddx);
            dvwdy = upw3(v_in_wm, b->ca[T_VW], b->ck[T_VW], j, W(i, j + 2, k), W(i, j + 1, k),
                         W(i, j, k), W(i, j - 1, k), W(i, j - 2, k), ddye, ddy);
            dwwdz = upw3(W(i, j, k), b->ca[T_WW], b->ck[T_WW], k, W(i, j, k + 2), W(i, j, k + 1),
                         W(i, j, k), W(i, j, k - 1), W(i, j, k - 2), V1(b->deltaz, k + 1), V1(b->deltaz, k));
        } else {
            duwdx = -0.5 * (u15 * w5 + alpha * fabs(u15) * w6 - u16 * w7 - alpha * fabs(u16) * w8) / V1(b->deltax, i);
            dvwdy = -0.5 * (v15 * w9 + alpha * fabs(v15) * w10 - v16 * w11 - alpha * fabs(v16) * w12) / V1(b->deltay, j);
            dwwdz = -0.5 * (w13 * w13 + alpha * fabs(w13) * w14 - w15 * w15 - alpha * fabs(w15) * w16) / V1(b->deltaz, k);
        }
        double d2wdx2 = (2.0 / (0.5 * (dx2xr + dx2xl))) * ((-w8 / (0.5 * dx2xr)) + (w6 / (0.5 * dx2xl)));
        double d2wdy2 = (2.0 / (0.5 * (dy2ye + dy2yw))) * ((-w12 / (0.5 * dy2ye)) + (w10 / (0.5 * dy2yw)));
        double d2wdz2 = (2.0 / dz2zt) * ((-w16 / dzt) + (w14 / dz));
        double ztt2 = (d2wdx2 + d2wdy2 + d2wdz2) / re;
        double residw = (-duwdx - dvwdy - dwwdz + ztt2);
        FLD(b, wt, i, j, k) = W(i, j, k) + dt * (residw + dpdz);

        /* ---- restore TS neighbours (NavierStokes.f90:860-907) ---- */
#define GT(q, idx) V1(b->ghost[q], idx)
        if (C2(i + 1, j, k) == 2) {
            U(i + 1, j, k) = GT(G_U2T, iip1); Vv(i + 1, j, k) = GT(G_V2T, iip1);
            Vv(i + 1, j - 1, k) = GT(G_V1T, iip1); W(i + 1, j, k) = GT(G_W2T, iip1);
            W(i + 1, j, k - 1) = GT(G_W1T, iip1); P(i + 1, j, k) = GT(G_PT, iip1);
        }
        if (C2(i - 1, j, k) == 2) {
            U(i - 2, j, k) = GT(G_U1T, iim1); Vv(i - 1, j, k) = GT(G_V2T, iim1);
            Vv(i - 1, j - 1, k) = GT(G_V1T, iim1); W(i - 1, j, k) = GT(G_W2T, iim1);
            W(i - 1, j, k - 1) = GT(G_W1T, iim1); P(i - 1, j, k) = GT(G_PT, iim1);
        }
        if (C2(i, j + 1, k) == 2) {
            U(i, j + 1, k) = GT(G_U2T, ijp1); U(i - 1, j + 1, k) = GT(G_U1T, ijp1);
            Vv(i, j + 1, k) = GT(G_V2T, ijp1); W(i, j + 1, k) = GT(G_W2T, ijp1);
            W(i, j + 1, k - 1) = GT(G_W1T, ijp1); P(i, j + 1, k) = GT(G_PT, ijp1);
        }
        if (C2(i, j - 1, k) == 2) {
            U(i, j - 1, k) = GT(G_U2T, ijm1); U(i - 1, j - 1, k) = GT(G_U1T, ijm1);
            Vv(i, j - 2, k) = GT(G_V1T, ijm1); W(i, j - 1, k) = GT(G_W2T, ijm1);
            W(i, j - 1, k - 1) = GT(G_W1T, ijm1); P(i, j - 1, k) = GT(G_PT, ijm1);
        }
        if (C2(i, j, k + 1) == 2) {
            U(i, j, k + 1) = GT(G_U2T, ikp1); U(i - 1, j, k + 1) = GT(G_U1T, ikp1);
            Vv(i, j, k + 1) = GT(G_V2T, ikp1); Vv(i, j - 1, k + 1) = GT(G_V1T, ikp1);
            W(i, j, k + 1) = GT(G_W2T, ikp1); P(i, j, k + 1) = GT(G_PT, ikp1);
        }
        if (C2(i, j, k - 1) == 2) {
            U(i, j, k - 1) = GT(G_U2T, ikm1); U(i - 1, j, k - 1) = GT(G_U1T, ikm1);
            Vv(i, j, k - 1) = GT(G_V2T, ikm1); Vv(i, j - 1, k - 1) = GT(G_V1T, ikm1);
            W(i, j, k - 2) = GT(G_W1T, ikm1); P(i, j, k - 1) = GT(G_PT, ikm1);
        }
#undef GT
    }
#undef U
#undef Vv
#undef W
#undef P
#undef C2
}
