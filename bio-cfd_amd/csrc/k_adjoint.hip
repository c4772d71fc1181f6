// Adjoint (transposed) operators of the pressure projection and of the
// coarse/fine trilinear updates, FP64, geometry frozen (tags, index sets and
// interface tables are the forward's).
//
// Projection of one block with a fixed SOR iteration count N (the forward is
// biocfd_projection: computeDiv, N red-black SOR iterations from pc = 0,
// correctPressure + correctVelocity, PcorVcor.F90:222-377):
//
//   b   = D (ut, vt, wt)                       (computeDiv, fluid cells)
//   pc  = S_N (b / dt)                         (R_1 B_1 ... R_N B_N from 0)
//   p'  = p + pc,  ut' = ut - G_x pc, ...      (fluid cells)
//
// The transpose runs backwards: lambda_pc = pbar + G^T (ubar, vbar, wbar) on
// the fluid cells, then B_N^T R_N^T ... B_1^T R_1^T, accumulating the
// adjoint of b, then utbar = ubar + D^T(bbar) and likewise for v, w.  A
// colour sweep of colour C updates x_y = (1 - w) x_y + w / D_y (rhs_y / dt -
// sum_n a_yn x_n) at the fluid cells y of colour C, reading only the other
// colour; its transpose, in gather form, is
//   (a) at every fluid cell x of the other colour:
//       l_x -= sum over fluid neighbours y of colour C of w a_{y->x} / D_y l_y
//       (a_{y->x}: the coefficient row y's update gives neighbour x, Acx(y,3)
//       for x east of y, Acx(y,1) for x west, and likewise in y and z);
//   (b) at every fluid cell y of colour C: bbar_y += w / (D_y dt) l_y,
//       l_y = (1 - w) l_y;
// (a) before (b).  Non-fluid cells hold pc = 0 in the forward (constants),
// so no adjoint flows through them.  Red means (i + j + k) odd in the
// reference's 1-based indices with the ghost layer.
//
// Coarse/fine updates (CoarseInterp.f90:447-502, FineInterp_bound.f90:554-603)
// overwrite each target with a trilinear combination of eight sources: the
// transpose adds w_s * l_target to each source (atomically: sources are
// shared between targets) and zeroes l_target, the prolongation's axis passes
// in reverse order (a cell two passes write keeps the last pass's value).
#include "biocfd_rt.hpp"

namespace bcfd {

namespace {

__device__ inline bool fluid(const Dims &d, const uint8_t *cell, int i, int j, int k) {
  return i >= 2 && i <= d.nx + 1 && j >= 2 && j <= d.ny + 1 && k >= 2 && k <= d.nz + 1 && cell[ix3(d, i, j, k)] == 0;
}

__device__ inline bool red(int i, int j, int k) { return ((i + j + k) & 1) == 1; }

// coefficient column m (0: west/south/bottom, 1: centre, 2: east/north/top)
// of row q (1-based cell index, interior) of an (n,3) table
__device__ inline double coef(const double *A, int n, int q, int m) { return A[(q - 2) + m * n]; }

__device__ inline double diag(const double *Ax, const double *Ay, const double *Az, const Dims &d, int i, int j,
                              int k) {
  return (coef(Ax, d.nx, i, 1) + coef(Ay, d.ny, j, 1)) + coef(Az, d.nz, k, 1);
}

struct Cell {
  int i, j, k;
  int64_t c;
};

__device__ inline bool cell_of(const Dims &d, Cell &x) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)d.NX * d.NY * d.NZ;
  if (q >= n) return false;
  x.c = q;
  x.i = 1 + (int)(q % d.NX);
  x.j = 1 + (int)((q / d.NX) % d.NY);
  x.k = 1 + (int)(q / ((int64_t)d.NX * d.NY));
  return true;
}

// lambda_pc = pbar + G^T (ubar, vbar, wbar) on fluid cells, 0 elsewhere;
// bbar = 0.  correctVelocity: ut(c) -= dt / (0.5 (dx(i+1) + dx(i))) (pc(c+1)
// - pc(c)) at fluid c (PcorVcor.F90:268-285, k_correct's expressions)
__global__ void k_adj_grad(Dims d, const uint8_t *__restrict__ cell, const double *__restrict__ ub,
                           const double *__restrict__ vb, const double *__restrict__ wb,
                           const double *__restrict__ pb, const double *__restrict__ dxv,
                           const double *__restrict__ dyv, const double *__restrict__ dzv, double dt,
                           double *__restrict__ lpc, double *__restrict__ bbar) {
  Cell x;
  if (!cell_of(d, x)) return;
  bbar[x.c] = 0.0;
  if (!fluid(d, cell, x.i, x.j, x.k)) {
    lpc[x.c] = 0.0;
    return;
  }
  const int i = x.i, j = x.j, k = x.k;
  const int64_t sy = d.NX, sz = (int64_t)d.NX * d.NY;
  double l = pb[x.c];
  l += dt / (0.5 * (dxv[i] + dxv[i - 1])) * ub[x.c];
  l += dt / (0.5 * (dyv[j] + dyv[j - 1])) * vb[x.c];
  l += dt / (0.5 * (dzv[k] + dzv[k - 1])) * wb[x.c];
  if (fluid(d, cell, i - 1, j, k)) l -= dt / (0.5 * (dxv[i - 1] + dxv[i - 2])) * ub[x.c - 1];
  if (fluid(d, cell, i, j - 1, k)) l -= dt / (0.5 * (dyv[j - 1] + dyv[j - 2])) * vb[x.c - sy];
  if (fluid(d, cell, i, j, k - 1)) l -= dt / (0.5 * (dzv[k - 1] + dzv[k - 2])) * wb[x.c - sz];
  lpc[x.c] = l;
}

// transpose of a colour-C sweep, step (a): fluid cells of the other colour
// gather from their fluid neighbours of colour C
__global__ void k_adj_sor_a(Dims d, const uint8_t *__restrict__ cell, const double *__restrict__ Ax,
                            const double *__restrict__ Ay, const double *__restrict__ Az, double omega,
                            int colour_c, double *__restrict__ lpc) {
  Cell x;
  if (!cell_of(d, x)) return;
  const int i = x.i, j = x.j, k = x.k;
  if ((int)red(i, j, k) == colour_c || !fluid(d, cell, i, j, k)) return;
  const int64_t sy = d.NX, sz = (int64_t)d.NX * d.NY;
  double acc = 0.0;
  // neighbour y west of x: x is y's east neighbour (column 2), and so on
  if (fluid(d, cell, i - 1, j, k))
    acc += omega * coef(Ax, d.nx, i - 1, 2) / diag(Ax, Ay, Az, d, i - 1, j, k) * lpc[x.c - 1];
  if (fluid(d, cell, i + 1, j, k))
    acc += omega * coef(Ax, d.nx, i + 1, 0) / diag(Ax, Ay, Az, d, i + 1, j, k) * lpc[x.c + 1];
  if (fluid(d, cell, i, j - 1, k))
    acc += omega * coef(Ay, d.ny, j - 1, 2) / diag(Ax, Ay, Az, d, i, j - 1, k) * lpc[x.c - sy];
  if (fluid(d, cell, i, j + 1, k))
    acc += omega * coef(Ay, d.ny, j + 1, 0) / diag(Ax, Ay, Az, d, i, j + 1, k) * lpc[x.c + sy];
  if (fluid(d, cell, i, j, k - 1))
    acc += omega * coef(Az, d.nz, k - 1, 2) / diag(Ax, Ay, Az, d, i, j, k - 1) * lpc[x.c - sz];
  if (fluid(d, cell, i, j, k + 1))
    acc += omega * coef(Az, d.nz, k + 1, 0) / diag(Ax, Ay, Az, d, i, j, k + 1) * lpc[x.c + sz];
  lpc[x.c] -= acc;
}

// step (b): fluid cells of colour C
__global__ void k_adj_sor_b(Dims d, const uint8_t *__restrict__ cell, const double *__restrict__ Ax,
                            const double *__restrict__ Ay, const double *__restrict__ Az, double omega, double dt,
                            int colour_c, double *__restrict__ lpc, double *__restrict__ bbar) {
  Cell x;
  if (!cell_of(d, x)) return;
  const int i = x.i, j = x.j, k = x.k;
  if ((int)red(i, j, k) != colour_c || !fluid(d, cell, i, j, k)) return;
  const double l = lpc[x.c];
  bbar[x.c] += omega / (diag(Ax, Ay, Az, d, i, j, k) * dt) * l;
  lpc[x.c] = (1.0 - omega) * l;
}

// utbar = ubar + D^T bbar; computeDiv: b(i) = (ut(i) - ut(i-1)) / dx(i) + ...
// at fluid cells (PcorVcor.F90:229-242, k_div's expressions)
__global__ void k_adj_div(Dims d, const uint8_t *__restrict__ cell, const double *__restrict__ bbar,
                          const double *__restrict__ ub, const double *__restrict__ vb,
                          const double *__restrict__ wb, const double *__restrict__ dxv,
                          const double *__restrict__ dyv, const double *__restrict__ dzv,
                          double *__restrict__ utb, double *__restrict__ vtb, double *__restrict__ wtb) {
  Cell x;
  if (!cell_of(d, x)) return;
  const int i = x.i, j = x.j, k = x.k;
  const int64_t sy = d.NX, sz = (int64_t)d.NX * d.NY;
  double u = ub[x.c], v = vb[x.c], w = wb[x.c];
  if (fluid(d, cell, i, j, k)) {
    u += bbar[x.c] / dxv[i - 1];
    v += bbar[x.c] / dyv[j - 1];
    w += bbar[x.c] / dzv[k - 1];
  }
  if (fluid(d, cell, i + 1, j, k)) u -= bbar[x.c + 1] / dxv[i];
  if (fluid(d, cell, i, j + 1, k)) v -= bbar[x.c + sy] / dyv[j];
  if (fluid(d, cell, i, j, k + 1)) w -= bbar[x.c + sz] / dzv[k];
  utb[x.c] = u;
  vtb[x.c] = v;
  wtb[x.c] = w;
}

// ---- coarse/fine interface transposes ----

struct ITabA {
  const int *det[3];
  int cnt[3];
};

__device__ inline int DETA(const ITabA &t, int ax, int row, int n) { return t.det[ax][(row - 1) + 3 * (n - 1)]; }

struct GridA {
  const double *x, *y, *z;
};

// bilinear weights of f1 .. f4 at (x, y) (the forward's bl: f1 (x2-x)(y2-y),
// f3 (x2-x)(y-y1), f2 (x-x1)(y2-y), f4 (x-x1)(y-y1), over (x2-x1)(y2-y1))
__device__ inline void bl_weights(double x, double y, double x1, double x2, double y1, double y2, double w[4]) {
  const double den = (x2 - x1) * (y2 - y1);
  w[0] = (x2 - x) * (y2 - y) / den;
  w[1] = (x - x1) * (y2 - y) / den;
  w[2] = (x2 - x) * (y - y1) / den;
  w[3] = (x - x1) * (y - y1) / den;
}

// transpose of k_restrict (one target coarse cell per thread, the same
// index arithmetic and guards)
__global__ void k_restrict_adj(Dims cd, Dims fd, ITabA t, int comp, int cintp, GridA cg, GridA fg,
                               double *__restrict__ lcoarse, double *__restrict__ lfine) {
  const int ni = t.cnt[0] - 2 * cintp + 1, nj = t.cnt[1] - 2 * cintp + 1, nk = t.cnt[2] - 2 * cintp + 1;
  if (ni <= 0 || nj <= 0 || nk <= 0) return;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)ni * nj * nk) return;
  const int i = cintp + (int)(q % ni), j = cintp + (int)((q / ni) % nj), k = cintp + (int)(q / ((int64_t)ni * nj));
  const int tx = DETA(t, 0, 1, i), ty = DETA(t, 1, 1, j), tz = DETA(t, 2, 1, k);
  int lx = (DETA(t, 0, 2, i) + DETA(t, 0, 3, i) - 1) / 2;
  int ly = (DETA(t, 1, 2, j) + DETA(t, 1, 3, j) - 1) / 2;
  int lz = (DETA(t, 2, 2, k) + DETA(t, 2, 3, k) - 1) / 2;
  int ox = 0, oy = 0, oz = 0;
  if (comp == 1) { lx = DETA(t, 0, 3, i); ox = 1; }
  if (comp == 2) { ly = DETA(t, 1, 3, j); oy = 1; }
  if (comp == 3) { lz = DETA(t, 2, 3, k); oz = 1; }
  int ii = lx, jj = ly, kk = lz, vi = lx, vj = ly, vk = lz;
  if (comp == 1) { ii = lx - 1; vi = lx - 2; }
  else if (comp == 2) { jj = ly - 1; vj = ly - 2; }
  else if (comp == 3) { kk = lz - 1; vk = lz - 2; }
  if (vi < 1 || vj < 1 || vk < 1 || vi + 1 > fd.NX || vj + 1 > fd.NY || vk + 1 > fd.NZ || ii < 1 || jj < 1 ||
      kk < 1 || tx - ox < 1 || ty - oy < 1 || tz - oz < 1 || tx > cd.NX + 1 || ty > cd.NY + 1 || tz > cd.NZ + 1)
    return;
  const double x = cg.x[tx - 1], y = cg.y[ty - 1], z = cg.z[tz - 1];
  double w[4];
  bl_weights(x, y, fg.x[ii - 1], fg.x[ii], fg.y[jj - 1], fg.y[jj], w);
  const double Z0 = fg.z[kk - 1], Z1 = fg.z[kk];
  const double tz_ = (z - Z0) / (Z1 - Z0);
  const int64_t dst = ix3(cd, tx - ox, ty - oy, tz - oz);
  const double l = lcoarse[dst];
  lcoarse[dst] = 0.0;
  const int64_t s0 = ix3(fd, vi, vj, vk), s1 = ix3(fd, vi + 1, vj, vk), s2 = ix3(fd, vi, vj + 1, vk),
                s3 = ix3(fd, vi + 1, vj + 1, vk), up = (int64_t)fd.NX * fd.NY;
  const int64_t src[4] = {s0, s1, s2, s3};
  for (int m = 0; m < 4; ++m) {
    atomicAdd(&lfine[src[m]], (1.0 - tz_) * w[m] * l);
    atomicAdd(&lfine[src[m] + up], tz_ * w[m] * l);
  }
}

// transpose of one axis pass of k_prolong
__global__ void k_prolong_adj(Dims cd, Dims fd, ITabA t, int comp, int axis, GridA cg, GridA fg,
                              double *__restrict__ lfine, double *__restrict__ lcoarse, int factor) {
  int n[3] = {t.cnt[0], t.cnt[1], t.cnt[2]};
  n[axis] = 2;
  const int64_t nent = (int64_t)n[0] * n[1] * n[2];
  const int per = factor * factor * factor;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nent * per) return;
  const int64_t e = q / per;
  const int sub = (int)(q % per);
  int idx[3] = {(int)(e % n[0]), (int)((e / n[0]) % n[1]), (int)(e / ((int64_t)n[0] * n[1]))};
  for (int a = 0; a < 3; ++a) idx[a] += 1;
  if (idx[axis] == 2) idx[axis] = t.cnt[axis];
  const int i = idx[0], j = idx[1], k = idx[2];
  const int x1 = DETA(t, 0, 2, i), x2 = DETA(t, 0, 3, i);
  const int y1 = DETA(t, 1, 2, j), y2 = DETA(t, 1, 3, j);
  const int z1 = DETA(t, 2, 2, k), z2 = DETA(t, 2, 3, k);
  const int sx = sub % factor, sy = (sub / factor) % factor, sz = sub / (factor * factor);
  const int tx = x1 + sx, ty = y1 + sy, tz = z1 + sz;
  if (tx > x2 || ty > y2 || tz > z2) return;
  const int lx = DETA(t, 0, 1, i), ly = DETA(t, 1, 1, j), lz = DETA(t, 2, 1, k);
  int vi = lx, vj = ly, vk = lz;
  if (comp == 1) vi = lx - 1;
  else if (comp == 2) vj = ly - 1;
  else if (comp == 3) vk = lz - 1;
  const int ox = (comp == 1), oy = (comp == 2), oz = (comp == 3);
  if (vi - 1 < 1 || vj - 1 < 1 || vk - 1 < 1 || vi + 1 > cd.NX || vj + 1 > cd.NY || vk + 1 > cd.NZ || lx < 2 ||
      ly < 2 || lz < 2 || tx - ox < 1 || ty - oy < 1 || tz - oz < 1 || tx - ox > fd.NX || ty - oy > fd.NY ||
      tz - oz > fd.NZ)
    return;
  const double x = fg.x[tx - 1], y = fg.y[ty - 1], z = fg.z[tz - 1];
  double w[4];
  bl_weights(x, y, cg.x[lx - 2], cg.x[lx], cg.y[ly - 2], cg.y[ly], w);
  const double Z0 = cg.z[lz - 2], Z1 = cg.z[lz];
  const double tz_ = (z - Z0) / (Z1 - Z0);
  const int64_t dst = ix3(fd, tx - ox, ty - oy, tz - oz);
  const double l = lfine[dst];
  lfine[dst] = 0.0;
  const int64_t s0 = ix3(cd, vi - 1, vj - 1, vk - 1), s1 = ix3(cd, vi + 1, vj - 1, vk - 1),
                s2 = ix3(cd, vi - 1, vj + 1, vk - 1), s3 = ix3(cd, vi + 1, vj + 1, vk - 1),
                up = 2 * (int64_t)cd.NX * cd.NY;
  const int64_t src[4] = {s0, s1, s2, s3};
  for (int m = 0; m < 4; ++m) {
    atomicAdd(&lcoarse[src[m]], (1.0 - tz_) * w[m] * l);
    atomicAdd(&lcoarse[src[m] + up], tz_ * w[m] * l);
  }
}

GridA grids_a(const Block &b, int comp) {
  switch (comp) {
    case 0: return {b.g[BG_XP], b.g[BG_YP], b.g[BG_ZP]};
    case 1: return {b.g[BG_XU], b.g[BG_YU], b.g[BG_ZU]};
    case 2: return {b.g[BG_XV], b.g[BG_YV], b.g[BG_ZV]};
    default: return {b.g[BG_XW], b.g[BG_YW], b.g[BG_ZW]};
  }
}

ITabA tab_a(const Interface &it, int comp) {
  ITabA t;
  for (int a = 0; a < 3; ++a) { t.det[a] = it.det[comp][a]; t.cnt[a] = (int)it.counter[comp][a]; }
  return t;
}

// device scratch of the adjoint entry points (grown as needed, kept)
double *g_adj = nullptr;
int64_t g_adj_cap = 0;

int adj_scratch(int64_t n) { return grow(&g_adj, &g_adj_cap, n); }

int upload(double *dst, const double *host, int64_t n) {
  BCFD_HIP(hipMemcpyAsync(dst, host, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, rt().stream));
  return 0;
}

int download(double *host, const double *src, int64_t n) {
  BCFD_HIP(hipMemcpyAsync(host, src, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, rt().stream));
  return 0;
}

Interface *get_interface(int itf, int comp) {
  if (itf < 1 || itf > kMaxInterfaces || !rt().itf[itf].live) {
    set_error("no interface %d", itf);
    return nullptr;
  }
  if (comp < 0 || comp > 3) {
    set_error("component %d outside 0 (p) .. 3 (w)", comp);
    return nullptr;
  }
  return &rt().itf[itf];
}

}  // namespace

}  // namespace bcfd

using namespace bcfd;

extern "C" {

int biocfd_projection_adjoint(int id, int64_t n_iter, double dt, const double *ubar, const double *vbar,
                              const double *wbar, const double *pbar, double *utbar, double *vtbar, double *wtbar) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (!b->A[0] || !b->g[BG_DX]) { set_error("projection_adjoint: block %d is not set up", id); return -1; }
  if (n_iter < 1) { set_error("projection_adjoint: iteration count %lld < 1", (long long)n_iter); return -1; }
  const int64_t n = b->n2;
  if (adj_scratch(7 * n)) return -1;
  double *ub = g_adj, *vb = ub + n, *wb = vb + n, *pb = wb + n, *lpc = pb + n, *bbar = lpc + n, *out = bbar + n;
  if (upload(ub, ubar, n) || upload(vb, vbar, n) || upload(wb, wbar, n) || upload(pb, pbar, n)) return -1;
  hipStream_t s = rt().stream;
  const dim3 blk(256), grd(div_up(n, 256));
  const Dims &d = b->d;
  const uint8_t *cell = b->tag[BT_CELL];
  hipLaunchKernelGGL(k_adj_grad, grd, blk, 0, s, d, cell, ub, vb, wb, pb, b->g[BG_DX], b->g[BG_DY], b->g[BG_DZ], dt,
                     lpc, bbar);
  for (int64_t it = 0; it < n_iter; ++it) {
    // black sweep transposed, then red (the forward's last sweep first)
    for (int colour : {0, 1}) {
      hipLaunchKernelGGL(k_adj_sor_a, grd, blk, 0, s, d, cell, b->A[0], b->A[1], b->A[2], b->omega, colour, lpc);
      hipLaunchKernelGGL(k_adj_sor_b, grd, blk, 0, s, d, cell, b->A[0], b->A[1], b->A[2], b->omega, dt, colour,
                         lpc, bbar);
    }
  }
  // utbar into out, vtbar into lpc, wtbar into pb (neither is read any more)
  hipLaunchKernelGGL(k_adj_div, grd, blk, 0, s, d, cell, bbar, ub, vb, wb, b->g[BG_DX], b->g[BG_DY], b->g[BG_DZ],
                     out, lpc, pb);
  BCFD_CHECK_LAUNCH();
  if (download(utbar, out, n) || download(vtbar, lpc, n) || download(wtbar, pb, n)) return -1;
  BCFD_HIP(hipStreamSynchronize(s));
  return 0;
}

int biocfd_coarse_update_adjoint(int itf, int comp, double *lam_coarse, double *lam_fine) {
  BCFD_API_LOCK();
  Interface *it = get_interface(itf, comp);
  if (!it) return -1;
  Block &a = rt().blk[it->a], &f = rt().blk[it->b];
  const int64_t nc = a.n2, nf = f.n2;
  if (adj_scratch(nc + nf)) return -1;
  double *lc = g_adj, *lf = g_adj + nc;
  if (upload(lc, lam_coarse, nc) || upload(lf, lam_fine, nf)) return -1;
  const ITabA t = tab_a(*it, comp);
  const int cintp = (int)a.cintp;
  const int64_t total = (int64_t)(t.cnt[0] - 2 * cintp + 1) * (t.cnt[1] - 2 * cintp + 1) * (t.cnt[2] - 2 * cintp + 1);
  if (total > 0)
    hipLaunchKernelGGL(k_restrict_adj, dim3(div_up(total, 256)), dim3(256), 0, rt().stream, a.d, f.d, t, comp, cintp,
                       grids_a(a, comp), grids_a(f, comp), lc, lf);
  BCFD_CHECK_LAUNCH();
  if (download(lam_coarse, lc, nc) || download(lam_fine, lf, nf)) return -1;
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  return 0;
}

int biocfd_fine_update_bd_adjoint(int itf, int comp, double *lam_fine, double *lam_coarse) {
  BCFD_API_LOCK();
  Interface *it = get_interface(itf, comp);
  if (!it) return -1;
  Block &a = rt().blk[it->a], &f = rt().blk[it->b];
  const int64_t nc = a.n2, nf = f.n2;
  if (adj_scratch(nc + nf)) return -1;
  double *lc = g_adj, *lf = g_adj + nc;
  if (upload(lc, lam_coarse, nc) || upload(lf, lam_fine, nf)) return -1;
  const ITabA t = tab_a(*it, comp);
  const int factor = (int)(it->b_msh / it->a_msh);
  for (int axis = 2; axis >= 0; --axis) {
    int n[3] = {t.cnt[0], t.cnt[1], t.cnt[2]};
    n[axis] = 2;
    const int64_t total = (int64_t)n[0] * n[1] * n[2] * factor * factor * factor;
    hipLaunchKernelGGL(k_prolong_adj, dim3(div_up(total, 256)), dim3(256), 0, rt().stream, a.d, f.d, t, comp, axis,
                       grids_a(a, comp), grids_a(f, comp), lf, lc, factor);
  }
  BCFD_CHECK_LAUNCH();
  if (download(lam_coarse, lc, nc) || download(lam_fine, lf, nf)) return -1;
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  return 0;
}

}  // extern "C"
