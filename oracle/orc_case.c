/*
 * Test infrastructure: case loading, set-up, time loop and record dumps for
 * the C oracle.  The call order restates the reference driver
 * app/main.F90:43-265 exactly as oracle/ref/ref_harness.F90 drives the
 * reference modules, so the two produce comparable dumps.
 */
#include <stdio.h>
#include <string.h>
#include <time.h>
#include "orc_internal.h"

void *orc_xcalloc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz ? sz : 1);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

double orc_noise_u(int64_t seed, int64_t field_id, int64_t lin)
{
    const int64_t pm = 2147483647LL;
    int64_t x = (seed + 2654435LL * field_id + lin) % pm;
    if (x < 0) x += pm;
    for (int r = 0; r < 3; ++r) x = (48271LL * x + 11LL) % pm;
    return (double)x / 2147483647.0;
}

typedef struct { const unsigned char *p; size_t n, pos; int err; } rd_t;

static int64_t rd_i8(rd_t *r)
{
    int64_t v = 0;
    if (r->pos + 8 > r->n) { r->err = 1; return 0; }
    memcpy(&v, r->p + r->pos, 8);
    r->pos += 8;
    return v;
}
static double rd_f8(rd_t *r)
{
    double v = 0;
    if (r->pos + 8 > r->n) { r->err = 1; return 0; }
    memcpy(&v, r->p + r->pos, 8);
    r->pos += 8;
    return v;
}
static double *rd_f8a(rd_t *r, int64_t n)
{
    double *a = orc_xcalloc((size_t)n, sizeof(double));
    for (int64_t i = 0; i < n; ++i) a[i] = rd_f8(r);
    return a;
}
static int64_t *rd_i8a(rd_t *r, int64_t n)
{
    int64_t *a = orc_xcalloc((size_t)n, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) a[i] = rd_i8(r);
    return a;
}

static void alloc_block_arrays(orc_block *b)
{
    size_t n2 = (size_t)(NX2(b) * NY2(b) * NZ2(b));
    size_t n3 = (size_t)((b->nx + 3) * (b->ny + 3) * (b->nz + 3));
    double **f[10] = { &b->u, &b->v, &b->w, &b->ut, &b->vt, &b->wt, &b->p, &b->b, &b->pc, &b->pco };
    for (int q = 0; q < 10; ++q) *f[q] = orc_xcalloc(n2, sizeof(double));
    b->cell = orc_xcalloc(n2, sizeof(int64_t));
    b->cell_n = orc_xcalloc(n2, sizeof(int64_t));
    b->cell2 = orc_xcalloc(n3, sizeof(int64_t));
    b->cell_pr = orc_xcalloc(n3, sizeof(int64_t));
    b->nodeIdTag = orc_xcalloc(n3, sizeof(int64_t));
    b->Acx = orc_xcalloc((size_t)(3 * b->nx), sizeof(double));
    b->Acy = orc_xcalloc((size_t)(3 * b->ny), sizeof(double));
    b->Acz = orc_xcalloc((size_t)(3 * b->nz), sizeof(double));
    size_t e = (size_t)(b->ibElems > 0 ? b->ibElems : 0);
    double **g[7] = { &b->xcent, &b->ycent, &b->zcent, &b->cosA, &b->cosB, &b->cosG, &b->elen };
    for (int q = 0; q < 7; ++q) *g[q] = orc_xcalloc(e, sizeof(double));
}

int orc_case_load(orc_case *c, const char *path)
{
    memset(c, 0, sizeof(*c));
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    fseek(fp, 0, SEEK_END);
    long sz = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    unsigned char *buf = orc_xcalloc((size_t)sz, 1);
    if (fread(buf, 1, (size_t)sz, fp) != (size_t)sz) { fclose(fp); free(buf); return -2; }
    fclose(fp);
    rd_t r = { buf, (size_t)sz, 0, 0 };
    int64_t magic = rd_i8(&r), version = rd_i8(&r);
    if (magic != 20261203 || version != 1) { free(buf); return -3; }
    c->nblocks = rd_i8(&r); c->nintf = rd_i8(&r); c->ib_block1 = rd_i8(&r);
    c->pcItaMax = rd_i8(&r); c->nsteps = rd_i8(&r); c->inor = rd_i8(&r); c->seed = rd_i8(&r);
    c->dt = rd_f8(&r); c->re = rd_f8(&r); c->uc = rd_f8(&r); c->epsi = rd_f8(&r);
    c->noise_amp = rd_f8(&r); c->aoa = rd_f8(&r); c->phase_angle_deg = rd_f8(&r);
    c->piv_pt = rd_f8(&r); c->dxmin = rd_f8(&r); c->freq = rd_f8(&r); c->mu_f = rd_f8(&r);
    c->rho_f = rd_f8(&r);
    int64_t nb = c->nblocks < 2 ? 2 : c->nblocks;
    c->blk = orc_xcalloc((size_t)nb, sizeof(orc_block));
    for (int64_t g = 0; g < c->nblocks; ++g) {
        orc_block *b = &c->blk[g];
        b->nx = rd_i8(&r); b->ny = rd_i8(&r); b->nz = rd_i8(&r); b->cintp = rd_i8(&r);
        b->has_ib = rd_i8(&r); b->moving = rd_i8(&r);
        b->dx = rd_f8(&r); b->dy = rd_f8(&r); b->dz = rd_f8(&r); b->omega = rd_f8(&r);
        int64_t nx = b->nx, ny = b->ny, nz = b->nz;
        b->x1 = rd_f8a(&r, nx + 3); b->y1 = rd_f8a(&r, ny + 3); b->z1 = rd_f8a(&r, nz + 3);
        b->deltax = rd_f8a(&r, nx + 2); b->deltay = rd_f8a(&r, ny + 2); b->deltaz = rd_f8a(&r, nz + 2);
        b->xu = rd_f8a(&r, nx + 3); b->yu = rd_f8a(&r, ny + 2); b->zu = rd_f8a(&r, nz + 2);
        b->xv = rd_f8a(&r, nx + 2); b->yv = rd_f8a(&r, ny + 3); b->zv = rd_f8a(&r, nz + 2);
        b->xw = rd_f8a(&r, nx + 2); b->yw = rd_f8a(&r, ny + 2); b->zw = rd_f8a(&r, nz + 3);
        b->xp = rd_f8a(&r, nx + 2); b->yp = rd_f8a(&r, ny + 2); b->zp = rd_f8a(&r, nz + 2);
        if (b->has_ib) {
            b->ibNodes = rd_i8(&r); b->ibElems = rd_i8(&r);
            for (int q = 0; q < 6; ++q) b->search[q] = rd_i8(&r);
            b->xshift = rd_f8(&r); b->yshift = rd_f8(&r); b->zshift = rd_f8(&r);
            b->a0 = rd_f8(&r); b->yamp = rd_f8(&r); b->bfreq = rd_f8(&r);
            b->xnode = rd_f8a(&r, b->ibNodes); b->ynode = rd_f8a(&r, b->ibNodes); b->znode = rd_f8a(&r, b->ibNodes);
            b->ibNodeId = rd_i8a(&r, b->ibNodes);
            b->elP1 = rd_i8a(&r, b->ibElems); b->elP2 = rd_i8a(&r, b->ibElems); b->elP3 = rd_i8a(&r, b->ibElems);
            b->surfId = rd_i8a(&r, b->ibElems);
        }
    }
    c->intf = orc_xcalloc((size_t)(c->nintf > 0 ? c->nintf : 1), sizeof(orc_interface));
    for (int64_t g = 0; g < c->nintf; ++g) {
        orc_interface *it = &c->intf[g];
        it->a_blk = rd_i8(&r); it->b_blk = rd_i8(&r); it->a_msh = rd_i8(&r); it->b_msh = rd_i8(&r);
        it->xs = rd_f8(&r); it->xe = rd_f8(&r); it->ys = rd_f8(&r); it->ye = rd_f8(&r);
        it->zs = rd_f8(&r); it->ze = rd_f8(&r);
    }
    c->ndump = rd_i8(&r);
    c->dump_steps = rd_i8a(&r, c->ndump);
    free(buf);
    return r.err ? -4 : 0;
}

static void free_block(orc_block *b)
{
    void *ptrs[] = { b->x1, b->y1, b->z1, b->deltax, b->deltay, b->deltaz, b->xu, b->yu, b->zu,
                     b->xv, b->yv, b->zv, b->xw, b->yw, b->zw, b->xp, b->yp, b->zp, b->u, b->v,
                     b->w, b->ut, b->vt, b->wt, b->p, b->b, b->pc, b->pco, b->cell, b->cell_n,
                     b->cell2, b->cell_pr, b->nodeIdTag, b->Acx, b->Acy, b->Acz, b->fluid,
                     b->red, b->black, b->solid, b->intercepted, b->ts, b->index_ts, b->nelp,
                     b->nelu1, b->nelu2, b->nelv1, b->nelv2, b->nelw1, b->nelw2, b->pNormDis,
                     b->u1NormDis, b->u2NormDis, b->v1NormDis, b->v2NormDis, b->w1NormDis,
                     b->w2NormDis, b->ibNodeId, b->elP1, b->elP2, b->elP3, b->surfId, b->xnode,
                     b->ynode, b->znode, b->xnode1, b->ynode1, b->znode1, b->xcent, b->ycent,
                     b->zcent, b->cosA, b->cosB, b->cosG, b->elen, b->u_dum, b->v_dum, b->w_dum,
                     b->p_dum };
    for (size_t q = 0; q < sizeof(ptrs) / sizeof(ptrs[0]); ++q) free(ptrs[q]);
    for (int q = 0; q < ORC_NGHOST; ++q) free(b->ghost[q]);
    for (int t = 0; t < T_COUNT; ++t) { free(b->ca[t]); free(b->ck[t]); }
}

void orc_case_free(orc_case *c)
{
    int64_t nb = c->nblocks < 2 ? 2 : c->nblocks;
    for (int64_t g = 0; g < nb && c->blk; ++g) free_block(&c->blk[g]);
    for (int64_t g = 0; g < c->nintf; ++g)
        for (int q = 0; q < 4; ++q)
            for (int a = 0; a < 3; ++a) free(c->intf[g].det[q][a]);
    free(c->blk); free(c->intf); free(c->dump_steps);
    memset(c, 0, sizeof(*c));
}

static void initial_conditions(orc_block *b, double uc)
{
    int64_t n2 = NX2(b) * NY2(b) * NZ2(b);
    for (int64_t q = 0; q < n2; ++q) {
        b->u[q] = uc; b->ut[q] = uc; b->v[q] = 0.0; b->vt[q] = 0.0;
        b->w[q] = 0.0; b->wt[q] = 0.0; b->p[q] = 0.0;
    }
    int64_t n3 = (b->nx + 3) * (b->ny + 3) * (b->nz + 3);
    for (int64_t q = 0; q < n3; ++q) b->cell_pr[q] = 0;
}

static void add_noise(orc_case *c, int64_t g)
{
    orc_block *b = &c->blk[g - 1];
    if (c->noise_amp == 0.0) return;
    int64_t s = c->seed + 1000 * g;
    for (int64_t n = 1; n <= b->fluidCount; ++n) {
        int64_t i = LIST(b->fluid, b->fluidCount, n, 1), j = LIST(b->fluid, b->fluidCount, n, 2),
                k = LIST(b->fluid, b->fluidCount, n, 3);
        int64_t lin = IX(b, i, j, k);
        FLD(b, u, i, j, k) = FLD(b, u, i, j, k) + c->noise_amp * (2.0 * orc_noise_u(s, 1, lin) - 1.0);
        FLD(b, v, i, j, k) = FLD(b, v, i, j, k) + c->noise_amp * (2.0 * orc_noise_u(s, 2, lin) - 1.0);
        FLD(b, w, i, j, k) = FLD(b, w, i, j, k) + c->noise_amp * (2.0 * orc_noise_u(s, 3, lin) - 1.0);
        FLD(b, ut, i, j, k) = FLD(b, u, i, j, k);
        FLD(b, vt, i, j, k) = FLD(b, v, i, j, k);
        FLD(b, wt, i, j, k) = FLD(b, w, i, j, k);
    }
}

static double g_phase;

int orc_case_setup(orc_case *c)
{
    int64_t nb = c->nblocks;
    if (c->nintf > 0) orc_interface_detail(c);
    for (int64_t g = 1; g <= nb; ++g) alloc_block_arrays(&c->blk[g - 1]);
    g_phase = c->phase_angle_deg * ORC_PI / 180.0;
    for (int64_t g = 1; g <= nb; ++g) {
        orc_block *b = &c->blk[g - 1];
        if (b->has_ib) {
            orc_find_dist_node(b);
            orc_shift_surface_nodes_initial(b, c->aoa, c->piv_pt, 0, c->dt, c->dxmin, 0.0);
            orc_compute_surface_norm(b, c->inor);
        }
    }
    for (int64_t g = 1; g <= nb; ++g) {
        orc_block *b = &c->blk[g - 1];
        if (b->has_ib) {
            orc_tagging_core(b);
            orc_cell_count_solid(b);
        }
    }
    if (!c->blk[0].has_ib) {
        orc_fine_block_cell(c);
        orc_cell_count_solid_coarse(&c->blk[0]);
        c->blk[0].tsCount = 0;
    }
    /* a fine block without a body: every interior cell fluid, the lists
     * built as for the coarse block */
    for (int64_t g = 2; g <= nb; ++g)
        if (!c->blk[g - 1].has_ib) {
            orc_cell_count_solid_coarse(&c->blk[g - 1]);
            c->blk[g - 1].tsCount = 0;
        }
    for (int64_t g = 1; g <= nb; ++g) {
        initial_conditions(&c->blk[g - 1], c->uc);
        add_noise(c, g);
    }
    for (int64_t g = 1; g <= nb; ++g) {
        orc_block *b = &c->blk[g - 1];
        if (b->has_ib) {
            orc_compute_norm_distance(b);
            orc_find_ts_cells(b);
        }
        orc_coefficient_matrix(b, g == 1);
    }
    for (int64_t g = 1; g <= nb; ++g) orc_non_uni_coeff(&c->blk[g - 1]);
    return 0;
}

int orc_case_step(orc_case *c, int64_t ita, double *totime, double *totalTime)
{
    int64_t nb = c->nblocks;
    *totime = *totime + c->dt;
    for (int64_t g = 1; g <= nb; ++g) {
        orc_block *b = &c->blk[g - 1];
        orc_ns_momentum(b, c->dt, c->re);
        if (g == 1) orc_velocity_bc(b, c->dt, c->uc);
        if (b->has_ib) {
            orc_solid_cell_bc(b);
            orc_velocity_forcing1(b);
        }
        if (g == 1) orc_velocity_bc(b, c->dt, c->uc);
    }
    orc_poisson_solver(c, ita, *totime, totalTime);
    for (int64_t g = 1; g <= nb; ++g)
        if (c->blk[g - 1].has_ib) orc_pressure_forcing1(&c->blk[g - 1]);
    for (int64_t g = 1; g <= nb; ++g) {
        orc_block *b = &c->blk[g - 1];
        if (b->has_ib) {
            b->blk_mv_tag = 0;
            orc_compute_surface_variables(b, g_phase, c->piv_pt, ita, c->dt, c->dxmin, c->freq, *totime);
        }
    }
    /* block move (main.F90:178-203): only where the fine blocks carry the
     * marker nodes the move check reads */
    int64_t flcnt = 0;
    if (c->nintf > 0 && orc_move_markers_defined(c)) orc_block_move_check(c, &flcnt);
    for (int64_t g = 2; g <= nb; ++g) orc_change_block_coords(&c->blk[g - 1]);
    for (int64_t q = 1; q <= c->nintf; ++q)
        orc_change_block_coords_interfaces(&c->intf[q - 1], &c->blk[c->intf[q - 1].b_blk - 1]);
    for (int64_t q = 1; q <= c->nintf; ++q)
        orc_change_block_interface(&c->intf[q - 1], &c->blk[c->intf[q - 1].b_blk - 1]);
    /* the coarse block's interface masks and lists (main.F90:195-200) */
    if (!c->blk[0].has_ib) {
        orc_fine_block_cell(c);
        orc_cell_count_solid_coarse_mv(c, &flcnt);
    }
    for (int64_t g = 1; g <= nb; ++g)
        if (c->blk[g - 1].has_ib) orc_compute_surface_norm(&c->blk[g - 1], c->inor);
    for (int64_t g = 1; g <= nb; ++g)
        if (c->blk[g - 1].has_ib) orc_tagging_move(c, g);
    for (int64_t g = 1; g <= nb; ++g) {
        orc_block *b = &c->blk[g - 1];
        if (!b->has_ib) continue;
        orc_selective_retagging(b);
        b->blk_mv_tag = 0;
        orc_cell_count_solid(b);
        orc_solid_cell_bc_move(b);
        orc_update_velocity_newv(b);
        b->move_check = 0;
        orc_compute_norm_distance(b);
        orc_find_ts_cells(b);
        orc_velocity_forcing_field(b);
        orc_pressure_forcing_field(b);
        orc_velocity_forcing_ghost(b);
        orc_pressure_forcing_ghost(b);
    }
    return 0;
}

/* ------------------------------ dumps ------------------------------ */
static void w_hdr(FILE *f, const char *name, int64_t kind, int64_t nd)
{
    char nm[24];
    memset(nm, ' ', sizeof(nm));
    memcpy(nm, name, strlen(name) < 24 ? strlen(name) : 24);
    fwrite(nm, 1, 24, f);
    fwrite(&kind, 8, 1, f);
    fwrite(&nd, 8, 1, f);
}
static void w_arr(FILE *f, const char *name, int kind, int nd, const int64_t *shape, const void *data)
{
    w_hdr(f, name, kind, nd);
    int64_t cnt = 1;
    for (int q = 0; q < nd; ++q) { fwrite(&shape[q], 8, 1, f); cnt *= shape[q]; }
    if (cnt > 0) fwrite(data, 8, (size_t)cnt, f);
}
static void w_f3(FILE *f, const orc_block *b, const char *nm, const double *a)
{
    int64_t s[3] = { NX2(b), NY2(b), NZ2(b) };
    w_arr(f, nm, 1, 3, s, a);
}
static void w_i3(FILE *f, const orc_block *b, const char *nm, const int64_t *a, int tall)
{
    int64_t s[3] = { NX2(b) + tall, NY2(b) + tall, NZ2(b) + tall };
    w_arr(f, nm, 2, 3, s, a);
}
static void w_i2(FILE *f, const char *nm, const int64_t *a, int64_t cnt)
{
    int64_t s[2] = { cnt, 3 };
    w_arr(f, nm, 2, 2, s, a);
}
static void w_f2(FILE *f, const char *nm, const double *a, int64_t r, int64_t cols)
{
    int64_t s[2] = { r, cols };
    w_arr(f, nm, 1, 2, s, a);
}
static void w_f1(FILE *f, const char *nm, const double *a, int64_t n) { w_arr(f, nm, 1, 1, &n, a); }
static void w_i1(FILE *f, const char *nm, const int64_t *a, int64_t n) { w_arr(f, nm, 2, 1, &n, a); }
static void w_f0(FILE *f, const char *nm, double v) { w_hdr(f, nm, 1, 0); fwrite(&v, 8, 1, f); }
static void w_i0(FILE *f, const char *nm, int64_t v) { w_hdr(f, nm, 2, 0); fwrite(&v, 8, 1, f); }

int orc_case_dump(const orc_case *c, const char *out_dir, int64_t step)
{
    static const char *tab_names[T_COUNT] = { "uu", "vv", "ww", "uv", "uw", "vu", "vw", "wu", "wv" };
    static const char *gnames[ORC_NGHOST] = { "p_ghost", "pt_ghost", "u1_ghost", "u1t_ghost", "u2_ghost",
                                              "u2t_ghost", "v1_ghost", "v1t_ghost", "v2_ghost", "v2t_ghost",
                                              "w1_ghost", "w1t_ghost", "w2_ghost", "w2t_ghost" };
    for (int64_t g = 1; g <= c->nblocks; ++g) {
        const orc_block *b = &c->blk[g - 1];
        char path[1024];
        snprintf(path, sizeof(path), "%s/dump_s%lld_b%lld.bin", out_dir, (long long)step, (long long)g);
        FILE *f = fopen(path, "wb");
        if (!f) return -1;
        w_f3(f, b, "u", b->u); w_f3(f, b, "v", b->v); w_f3(f, b, "w", b->w); w_f3(f, b, "p", b->p);
        w_f3(f, b, "ut", b->ut); w_f3(f, b, "vt", b->vt); w_f3(f, b, "wt", b->wt);
        w_f3(f, b, "pc", b->pc); w_f3(f, b, "pco", b->pco); w_f3(f, b, "b", b->b);
        w_f1(f, "x1", b->x1, b->nx + 3); w_f1(f, "y1", b->y1, b->ny + 3); w_f1(f, "z1", b->z1, b->nz + 3);
        w_i3(f, b, "cell", b->cell, 0); w_i3(f, b, "cell2", b->cell2, 1);
        w_i3(f, b, "nodeIdTag", b->nodeIdTag, 1); w_i3(f, b, "cell_n", b->cell_n, 0);
        w_i3(f, b, "cell_pr", b->cell_pr, 1);
        w_i0(f, "nIterPcor", b->nIterPcor); w_f0(f, "derr2", b->derr2); w_f0(f, "derrStdSt", b->derrStdSt);
        w_i0(f, "fluidCellCount", b->fluidCount); w_i0(f, "solidCellCount", b->solidCount);
        w_i0(f, "ibCellCount", b->ibCount); w_i0(f, "redCellCount", b->redCount);
        w_i0(f, "blackCellCount", b->blackCount); w_i0(f, "TSCellCount", b->tsCount);
        w_i2(f, "fluidIndexPtr", b->fluid, b->fluidCount);
        w_i2(f, "redCellIndexPtr", b->red, b->redCount);
        w_i2(f, "blackCellIndexPtr", b->black, b->blackCount);
        w_f2(f, "Acx", b->Acx, b->nx, 3); w_f2(f, "Acy", b->Acy, b->ny, 3); w_f2(f, "Acz", b->Acz, b->nz, 3);
        int64_t ns[T_COUNT] = { b->nx, b->ny, b->nz, b->nx, b->nx, b->ny, b->ny, b->nz, b->nz };
        for (int t = 0; t < T_COUNT; ++t) {
            char nm[16];
            snprintf(nm, sizeof(nm), "ca_%s", tab_names[t]);
            if (b->ca[t]) w_f2(f, nm, b->ca[t], 6, ns[t] + 2);
            snprintf(nm, sizeof(nm), "ck_%s", tab_names[t]);
            if (b->ck[t]) w_f2(f, nm, b->ck[t], 6, ns[t] + 2);
        }
        if (b->has_ib) {
            w_i2(f, "interceptedIndexPtr", b->intercepted, b->ibCount);
            w_i2(f, "solidIndexPtr", b->solid, b->solidCount);
            w_i2(f, "TSIndexPtr", b->ts, b->tsCount);
            w_i1(f, "index_ts", b->index_ts, b->tsCount);
            w_i1(f, "nelp", b->nelp, b->ibCount); w_i1(f, "nelu1", b->nelu1, b->ibCount);
            w_i1(f, "nelu2", b->nelu2, b->ibCount); w_i1(f, "nelv1", b->nelv1, b->ibCount);
            w_i1(f, "nelv2", b->nelv2, b->ibCount); w_i1(f, "nelw1", b->nelw1, b->ibCount);
            w_i1(f, "nelw2", b->nelw2, b->ibCount);
            w_f1(f, "pNormDis", b->pNormDis, b->ibCount); w_f1(f, "u1NormDis", b->u1NormDis, b->ibCount);
            w_f1(f, "u2NormDis", b->u2NormDis, b->ibCount); w_f1(f, "v1NormDis", b->v1NormDis, b->ibCount);
            w_f1(f, "v2NormDis", b->v2NormDis, b->ibCount); w_f1(f, "w1NormDis", b->w1NormDis, b->ibCount);
            w_f1(f, "w2NormDis", b->w2NormDis, b->ibCount);
            w_f1(f, "xcent", b->xcent, b->ibElems); w_f1(f, "ycent", b->ycent, b->ibElems);
            w_f1(f, "zcent", b->zcent, b->ibElems); w_f1(f, "cosAlpha", b->cosA, b->ibElems);
            w_f1(f, "cosBeta", b->cosB, b->ibElems); w_f1(f, "cosGamma", b->cosG, b->ibElems);
            w_f1(f, "element_length", b->elen, b->ibElems);
            w_f1(f, "xnode1", b->xnode1, b->ibNodes); w_f1(f, "ynode1", b->ynode1, b->ibNodes);
            w_f1(f, "znode1", b->znode1, b->ibNodes);
            for (int q = 0; q < ORC_NGHOST; ++q) w_f1(f, gnames[q], b->ghost[q], b->tsCount);
            w_f0(f, "xdot", b->xdot); w_f0(f, "ydot", b->ydot); w_f0(f, "yddot", b->yddot);
            w_f0(f, "thetaDot1", b->thetaDot1); w_f0(f, "thetaDDot1", b->thetaDDot1);
            w_f0(f, "piv_x", b->piv_x); w_f0(f, "piv_y", b->piv_y); w_f0(f, "piv_z", b->piv_z);
            w_f0(f, "xpth2", b->xpth2); w_f0(f, "ymove", b->ymove);
            orc_stress_calculation((orc_block *)b, c->mu_f, c->rho_f);
            w_f0(f, "viscous_drag", b->viscous_drag); w_f0(f, "pressure_drag", b->pressure_drag);
            w_f0(f, "viscous_lift", b->viscous_lift); w_f0(f, "pressure_lift", b->pressure_lift);
        }
        fclose(f);
    }
    return 0;
}

static int wants_dump(const orc_case *c, int64_t step)
{
    for (int64_t q = 0; q < c->ndump; ++q)
        if (c->dump_steps[q] == step) return 1;
    return 0;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int orc_run_case(const char *case_path, const char *out_dir, int64_t nsteps_override, double *seconds)
{
    orc_case c;
    int rc = orc_case_load(&c, case_path);
    if (rc) return rc;
    if (nsteps_override > 0) c.nsteps = nsteps_override;
    orc_case_setup(&c);
    double totime = c.dt, totalTime = 0.0;
    if (!seconds && wants_dump(&c, 0)) orc_case_dump(&c, out_dir, 0);
    double t0 = now_s();
    for (int64_t ita = 1; ita <= c.nsteps; ++ita) {
        orc_case_step(&c, ita, &totime, &totalTime);
        if (!seconds && wants_dump(&c, ita)) orc_case_dump(&c, out_dir, ita);
    }
    double t1 = now_s();
    if (seconds) {
        *seconds = t1 - t0;
        if (out_dir && out_dir[0]) orc_case_dump(&c, out_dir, c.nsteps);
    }
    orc_case_free(&c);
    return 0;
}
