/*
 * Test infrastructure: FP64 restatement of the pressure-correction projection,
 * reference src/PcorVcor.F90 (poissonSolver :22-220, computeDiv :222-243,
 * correctPressure :245-259, correctVelocity :261-286, REDBLACKSOR_linear
 * :288-377, updateVelocity_newv :379-397) and src/BoundaryConditions.f90
 * (velocityBC :11-74, solidCellBC :76-92, solidCellBC_move :94-115).
 */
#include "orc_internal.h"

void orc_velocity_bc(orc_block *b, double dt, double uc)
{
    int64_t nx = b->nx, ny = b->ny, nz = b->nz;
    for (int64_t k = 2; k <= nz + 1; ++k)
        for (int64_t j = 2; j <= ny + 1; ++j) {
            FLD(b, ut, 1, j, k) = uc;
            FLD(b, vt, 1, j, k) = -FLD(b, vt, nx + 1, j, k);
            FLD(b, wt, 1, j, k) = -FLD(b, wt, nx + 1, j, k);
            FLD(b, ut, nx + 1, j, k) = FLD(b, u, nx + 1, j, k) -
                (dt / V1(b->deltax, nx + 1)) * uc * (FLD(b, u, nx + 1, j, k) - FLD(b, u, nx, j, k));
            FLD(b, vt, nx + 2, j, k) = -FLD(b, vt, nx + 1, j, k) + FLD(b, v, nx + 2, j, k) +
                FLD(b, v, nx + 1, j, k) -
                (2.0 * dt / V1(b->deltax, nx + 2)) * uc * (FLD(b, v, nx + 2, j, k) - FLD(b, v, nx + 1, j, k));
            FLD(b, wt, nx + 2, j, k) = -FLD(b, wt, nx + 1, j, k) + FLD(b, w, nx + 2, j, k) +
                FLD(b, w, nx + 1, j, k) -
                (2.0 * dt / V1(b->deltax, nx + 2)) * uc * (FLD(b, w, nx + 2, j, k) - FLD(b, w, nx + 1, j, k));
        }
    for (int64_t k = 2; k <= nz + 1; ++k)
        for (int64_t i = 2; i <= nx + 1; ++i) {
            FLD(b, ut, i, 1, k) = FLD(b, ut, i, 2, k);
            FLD(b, wt, i, 1, k) = FLD(b, wt, i, 2, k);
            FLD(b, vt, i, 2, k) = 0.0;
            FLD(b, ut, i, ny + 2, k) = FLD(b, ut, i, ny + 1, k);
            FLD(b, wt, i, ny + 2, k) = FLD(b, wt, i, ny + 1, k);
            FLD(b, vt, i, ny + 1, k) = 0.0;
        }
    for (int64_t j = 2; j <= ny + 1; ++j)
        for (int64_t i = 2; i <= nx + 1; ++i) {
            FLD(b, ut, i, j, 1) = FLD(b, ut, i, j, 2);
            FLD(b, vt, i, j, 1) = FLD(b, vt, i, j, 2);
            FLD(b, wt, i, j, 1) = 0.0;
            FLD(b, ut, i, j, nz + 2) = FLD(b, ut, i, j, nz + 1);
            FLD(b, vt, i, j, nz + 2) = FLD(b, vt, i, j, nz + 1);
            FLD(b, wt, i, j, nz + 1) = 0.0;
        }
}

void orc_solid_cell_bc(orc_block *b)
{
    for (int64_t n = 1; n <= b->solidCount; ++n) {
        int64_t i = LIST(b->solid, b->solidCount, n, 1);
        int64_t j = LIST(b->solid, b->solidCount, n, 2);
        int64_t k = LIST(b->solid, b->solidCount, n, 3);
        FLD(b, ut, i, j, k) = 0.0;
        FLD(b, vt, i, j, k) = 0.0;
        FLD(b, wt, i, j, k) = 0.0;
        FLD(b, p, i, j, k) = 0.0;
    }
}

void orc_solid_cell_bc_move(orc_block *b)
{
    if (b->move_check != 1) return;
    for (int64_t n = 1; n <= b->solidCount; ++n) {
        int64_t i = LIST(b->solid, b->solidCount, n, 1);
        int64_t j = LIST(b->solid, b->solidCount, n, 2);
        int64_t k = LIST(b->solid, b->solidCount, n, 3);
        FLD(b, ut, i, j, k) = 0.0; FLD(b, vt, i, j, k) = 0.0; FLD(b, wt, i, j, k) = 0.0;
        FLD(b, p, i, j, k) = 0.0;
        FLD(b, u, i, j, k) = 0.0; FLD(b, v, i, j, k) = 0.0; FLD(b, w, i, j, k) = 0.0;
    }
}

void orc_update_velocity_newv(orc_block *b)
{
    if (b->move_check != 1) return;
    int64_t n = NX2(b) * NY2(b) * NZ2(b);
    for (int64_t q = 0; q < n; ++q) { b->ut[q] = b->u[q]; b->vt[q] = b->v[q]; b->wt[q] = b->w[q]; }
}

void orc_compute_div(orc_block *b)
{
    for (int64_t n = 1; n <= b->fluidCount; ++n) {
        int64_t i = LIST(b->fluid, b->fluidCount, n, 1);
        int64_t j = LIST(b->fluid, b->fluidCount, n, 2);
        int64_t k = LIST(b->fluid, b->fluidCount, n, 3);
        FLD(b, b, i, j, k) = (FLD(b, ut, i, j, k) - FLD(b, ut, i - 1, j, k)) / V1(b->deltax, i) +
                             (FLD(b, vt, i, j, k) - FLD(b, vt, i, j - 1, k)) / V1(b->deltay, j) +
                             (FLD(b, wt, i, j, k) - FLD(b, wt, i, j, k - 1)) / V1(b->deltaz, k);
    }
}

void orc_correct_pressure(orc_block *b)
{
    for (int64_t n = 1; n <= b->fluidCount; ++n) {
        int64_t i = LIST(b->fluid, b->fluidCount, n, 1);
        int64_t j = LIST(b->fluid, b->fluidCount, n, 2);
        int64_t k = LIST(b->fluid, b->fluidCount, n, 3);
        FLD(b, p, i, j, k) = FLD(b, p, i, j, k) + FLD(b, pc, i, j, k);
    }
}

void orc_correct_velocity(orc_block *b, double dt)
{
    for (int64_t n = 1; n <= b->fluidCount; ++n) {
        int64_t i = LIST(b->fluid, b->fluidCount, n, 1);
        int64_t j = LIST(b->fluid, b->fluidCount, n, 2);
        int64_t k = LIST(b->fluid, b->fluidCount, n, 3);
        FLD(b, ut, i, j, k) = FLD(b, ut, i, j, k) -
            dt / (0.5 * (V1(b->deltax, i + 1) + V1(b->deltax, i))) * (FLD(b, pc, i + 1, j, k) - FLD(b, pc, i, j, k));
        FLD(b, vt, i, j, k) = FLD(b, vt, i, j, k) -
            dt / (0.5 * (V1(b->deltay, j + 1) + V1(b->deltay, j))) * (FLD(b, pc, i, j + 1, k) - FLD(b, pc, i, j, k));
        FLD(b, wt, i, j, k) = FLD(b, wt, i, j, k) -
            dt / (0.5 * (V1(b->deltaz, k + 1) + V1(b->deltaz, k))) * (FLD(b, pc, i, j, k + 1) - FLD(b, pc, i, j, k));
    }
}

static void sor_colour(orc_block *b, const int64_t *lst, int64_t cnt, const double *src,
                       double dt, double omega)
{
    int64_t nx = b->nx, ny = b->ny, nz = b->nz;
    for (int64_t n = 1; n <= cnt; ++n) {
        int64_t i = LIST(lst, cnt, n, 1), j = LIST(lst, cnt, n, 2), k = LIST(lst, cnt, n, 3);
        double gs = (FLD(b, b, i, j, k) / dt
                     - AC(b->Acx, nx, i - 1, 1) * src[IX(b, i - 1, j, k)]
                     - AC(b->Acx, nx, i - 1, 3) * src[IX(b, i + 1, j, k)]
                     - AC(b->Acy, ny, j - 1, 1) * src[IX(b, i, j - 1, k)]
                     - AC(b->Acy, ny, j - 1, 3) * src[IX(b, i, j + 1, k)]
                     - AC(b->Acz, nz, k - 1, 1) * src[IX(b, i, j, k - 1)]
                     - AC(b->Acz, nz, k - 1, 3) * src[IX(b, i, j, k + 1)]) /
                    (AC(b->Acx, nx, i - 1, 2) + AC(b->Acy, ny, j - 1, 2) + AC(b->Acz, nz, k - 1, 2));
        FLD(b, pc, i, j, k) = gs;
        FLD(b, pc, i, j, k) = (1.0 - omega) * FLD(b, pco, i, j, k) + omega * FLD(b, pc, i, j, k);
    }
}

void orc_redblack_sor(orc_block *b, int64_t pcItaMax, double dt, double epsi)
{
    double omega = b->omega;
    double derr4 = 11111.0;
    int64_t ntot = NX2(b) * NY2(b) * NZ2(b);
    b->derr2 = 0.0;
    for (;;) {
        b->nIterPcor += 1;
        sor_colour(b, b->red, b->redCount, b->pco, dt, omega);
        sor_colour(b, b->black, b->blackCount, b->pc, dt, omega);
        if (b->nIterPcor % 5 == 0) {
            derr4 = 0.0;
            for (int64_t n = 1; n <= b->fluidCount; ++n) {
                int64_t i = LIST(b->fluid, b->fluidCount, n, 1);
                int64_t j = LIST(b->fluid, b->fluidCount, n, 2);
                int64_t k = LIST(b->fluid, b->fluidCount, n, 3);
                double var = fabs(FLD(b, pc, i, j, k) - FLD(b, pco, i, j, k));
                derr4 = derr4 > var ? derr4 : var;
            }
        }
        for (int64_t q = 0; q < ntot; ++q) b->pco[q] = b->pc[q];
        b->derr2 = derr4;
        if (derr4 < epsi || b->nIterPcor > pcItaMax) break;
    }
}

void orc_poisson_solver(orc_case *c, int64_t ita, double totime, double *totalTime)
{
    (void)ita;
    int64_t nb = c->nblocks;
    for (int64_t g = 1; g <= nb; ++g) {
        orc_block *b = &c->blk[g - 1];
        int64_t ntot = NX2(b) * NY2(b) * NZ2(b);
        for (int64_t q = 0; q < ntot; ++q) { b->b[q] = 0.0; b->pc[q] = 0.0; b->pco[q] = 0.0; }
        b->nIterPcor = 0;
        b->derr2 = 0.0;
        b->derrStdSt = 0.0;
    }
    orc_fine_update_newv_bd(c);
    orc_coarse_update_newv(c);
    for (int64_t g = 1; g <= nb; ++g) {
        orc_compute_div(&c->blk[g - 1]);
        if (g != 1) orc_redblack_sor(&c->blk[g - 1], c->pcItaMax, c->dt, c->epsi);
    }
    orc_coarse_update_pc(c);
    orc_redblack_sor(&c->blk[0], c->pcItaMax, c->dt, c->epsi);
    orc_fine_update_pc_bd(c);
    for (int64_t g = 2; g <= nb; ++g) orc_redblack_sor(&c->blk[g - 1], c->pcItaMax, c->dt, c->epsi);
    orc_coarse_update_pc(c);
    for (int64_t g = 1; g <= nb; ++g) {
        orc_correct_pressure(&c->blk[g - 1]);
        orc_correct_velocity(&c->blk[g - 1], c->dt);
    }
    orc_velocity_bc(&c->blk[0], c->dt, c->uc);
    for (int64_t g = 1; g <= nb; ++g) {
        orc_block *b = &c->blk[g - 1];
        double err = 0.0;
        for (int64_t n = 1; n <= b->fluidCount; ++n) {
            int64_t i = LIST(b->fluid, b->fluidCount, n, 1);
            int64_t j = LIST(b->fluid, b->fluidCount, n, 2);
            int64_t k = LIST(b->fluid, b->fluidCount, n, 3);
            double eu = fabs((FLD(b, ut, i, j, k) - FLD(b, u, i, j, k))) / c->dt;
            double ev = fabs((FLD(b, vt, i, j, k) - FLD(b, v, i, j, k))) / c->dt;
            double ew = fabs((FLD(b, wt, i, j, k) - FLD(b, w, i, j, k))) / c->dt;
            double m = err;
            if (eu > m) m = eu;
            if (ev > m) m = ev;
            if (ew > m) m = ew;
            err = m;
        }
        b->derrStdSt = err;
    }
    for (int64_t g = 1; g <= nb; ++g) *totalTime = totime + *totalTime;
    for (int64_t g = 1; g <= nb; ++g) {
        orc_block *b = &c->blk[g - 1];
        int64_t ntot = NX2(b) * NY2(b) * NZ2(b);
        for (int64_t q = 0; q < ntot; ++q) { b->u[q] = b->ut[q]; b->v[q] = b->vt[q]; b->w[q] = b->wt[q]; }
    }
    for (int64_t g = 1; g <= c->nintf; ++g) orc_fine_update_bd_mv(c, g);
    orc_coarse_update(c);
}
