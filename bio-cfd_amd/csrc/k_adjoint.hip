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
#include <cstring>

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

// wd = w / D at fluid cells, 0 elsewhere (the ghost ring included), and the
// iterate moved to m = wd * lambda (m = 0 off the fluid cells)
__global__ void k_adj_wd(Dims d, const uint8_t *__restrict__ cell, const double *__restrict__ Ax,
                         const double *__restrict__ Ay, const double *__restrict__ Az, double omega,
                         double *__restrict__ wd, double *__restrict__ lpc) {
  Cell x;
  if (!cell_of(d, x)) return;
  const double w = fluid(d, cell, x.i, x.j, x.k) ? omega / diag(Ax, Ay, Az, d, x.i, x.j, x.k) : 0.0;
  wd[x.c] = w;
  lpc[x.c] = w * lpc[x.c];
}

// The transpose of one iteration (black sweep B after red sweep R) is
// B^T(a), B^T(b), R^T(a), R^T(b), and step (b) of one colour and step (a)
// of the other touch the same cells: one launch per colour does, at its
// fluid cells, the pending (b) (unless first) and then the gather (a) from the
// other colour's cells, which this launch does not write (unless last).
// n iterations are 2 n + 1 launches, red first and last.  In m = wd * lambda
// the steps read only m of the neighbours:
//   (b)  bbar_y += m_y / dt,  m_y = (1 - w) m_y
//   (a)  m_x -= wd_x * sum over neighbours y of a_{y->x} m_y
// (m = 0 at non-fluid cells).  One thread per interior cell of the launch's
// colour on a 3-D grid (64 x 4 threads: i of the colour, j; one plane per
// grid layer): thread (ii, j, k) takes i = 2 + 2 ii + p, p the parity that
// gives (i + j + k) odd for red, even for black.  A fluid cell's neighbours
// lie inside the ghost ring; the coefficient tables are read only for
// interior rows.
__global__ void __launch_bounds__(256) k_adj_sor_colour(Dims d, const uint8_t *__restrict__ cell,
                                                        const double *__restrict__ Ax, const double *__restrict__ Ay,
                                                        const double *__restrict__ Az, const double *__restrict__ wd,
                                                        double rdt, double omega, int red_cells, int first, int last,
                                                        double *__restrict__ m, double *__restrict__ bbar) {
  // two planes per thread (k and k + 1: independent cells, their loads in
  // flight together); j is wave-uniform (one row of 64 threads per wave)
  const int ii = blockIdx.x * 64 + threadIdx.x;
  const int j = 2 + blockIdx.y * 4 + __builtin_amdgcn_readfirstlane(threadIdx.y);
  if (j > d.ny + 1) return;
  const int sy = d.NX, sz = d.NX * d.NY;
  const double ays = j > 2 ? coef(Ay, d.ny, j - 1, 2) : 0.0, ayn = j <= d.ny ? coef(Ay, d.ny, j + 1, 0) : 0.0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int k = 2 + 2 * blockIdx.z + h;
    if (k > d.nz + 1) break;
    const int i = 2 + 2 * ii + ((red_cells + j + k) & 1);
    if (i > d.nx + 1) continue;
    const int c = (i - 1) + sy * (j - 1) + sz * (k - 1);
    if (cell[c] != 0) continue;
    double mc = m[c];
    if (!first) {
      bbar[c] += mc * rdt;
      mc = (1.0 - omega) * mc;
    }
    if (!last) {
      // neighbour y west of x: x is y's east neighbour (column 2), and so on
      double acc = 0.0;
      if (i > 2) acc += coef(Ax, d.nx, i - 1, 2) * m[c - 1];
      if (i <= d.nx) acc += coef(Ax, d.nx, i + 1, 0) * m[c + 1];
      acc += ays * m[c - sy];
      acc += ayn * m[c + sy];
      if (k > 2) acc += coef(Az, d.nz, k - 1, 2) * m[c - sz];
      if (k <= d.nz) acc += coef(Az, d.nz, k + 1, 0) * m[c + sz];
      mc -= wd[c] * acc;
    }
    m[c] = mc;
  }
}

// lpc <- (B_1^T R_1^T ... B_N^T R_N^T applied backwards), bbar accumulated
// (lpc ends as wd * lambda: only bbar is read afterwards)
void adj_sor_launches(const Block &b, int64_t n_iter, double dt, double *wd, double *lpc, double *bbar) {
  hipStream_t s = rt().stream;
  const dim3 blk(256), grd(div_up(b.n2, 256));
  const uint8_t *cell = b.tag[BT_CELL];
  hipLaunchKernelGGL(k_adj_wd, grd, blk, 0, s, b.d, cell, b.A[0], b.A[1], b.A[2], b.omega, wd, lpc);
  const dim3 bc(64, 4, 1),
      gc((unsigned)div_up((b.d.nx + 1) / 2, 64), (unsigned)div_up(b.d.ny, 4), (unsigned)div_up(b.d.nz, 2));
  for (int64_t q = 0; q <= 2 * n_iter; ++q)
    hipLaunchKernelGGL(k_adj_sor_colour, gc, bc, 0, s, b.d, cell, b.A[0], b.A[1], b.A[2], wd, 1.0 / dt, b.omega,
                       q % 2 == 0 ? 1 : 0, q == 0 ? 1 : 0, q == 2 * n_iter ? 1 : 0, lpc, bbar);
}

// the same launches; on the block's own adjoint buffers (w != nullptr) and
// with SOR graphs on, as one graph captured on first use and replayed while
// the iteration count, dt and the tag array stay the same
int adj_sor(Block &b, int64_t n_iter, double dt, double *wd, double *lpc, double *bbar, AdWork *w) {
  if (b.n2 >= ((int64_t)1 << 31)) {
    set_error("adjoint SOR: block of %lld cells (32-bit cell offsets)", (long long)b.n2);
    return -1;
  }
  if (!w || !sor_graphs_on()) {
    adj_sor_launches(b, n_iter, dt, wd, lpc, bbar);
    BCFD_CHECK_LAUNCH();
    return 0;
  }
  long long dt_bits;
  memcpy(&dt_bits, &dt, sizeof dt_bits);
  const long long key[3] = {(long long)n_iter, dt_bits, (long long)(intptr_t)b.tag[BT_CELL]};
  if (w->sor_exec && memcmp(key, w->sor_key, sizeof key)) {
    (void)hipGraphExecDestroy(w->sor_exec);
    w->sor_exec = nullptr;
  }
  hipStream_t s = rt().stream;
  if (!w->sor_exec) {
    BCFD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    adj_sor_launches(b, n_iter, dt, wd, lpc, bbar);
    const hipError_t el = hipGetLastError();
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    if (el != hipSuccess || ec != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      set_error("adjoint SOR graph capture: %s", hipGetErrorString(el != hipSuccess ? el : ec));
      return -1;
    }
    const hipError_t ei = hipGraphInstantiate(&w->sor_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    BCFD_HIP(ei);
    memcpy(w->sor_key, key, sizeof key);
  }
  BCFD_HIP(hipGraphLaunch(w->sor_exec, s));
  return 0;
}

// utbar = ubar + D^T bbar; computeDiv: b(i) = (ut(i) - ut(i-1)) / dx(i) + ...
// at fluid cells (PcorVcor.F90:229-242, k_div's expressions)
// (ub and utb may be the same array: each thread reads and writes only its
// own cell of them)
__global__ void k_adj_div(Dims d, const uint8_t *__restrict__ cell, const double *__restrict__ bbar,
                          const double *ub, const double *vb, const double *wb, const double *__restrict__ dxv,
                          const double *__restrict__ dyv, const double *__restrict__ dzv, double *utb, double *vtb,
                          double *wtb) {
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
  if (adj_sor(*b, n_iter, dt, out, lpc, bbar, nullptr)) return -1;
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

// ---------------------------------------------------------------------------
// The adjoint time step (no counterpart in the reference): sensitivity
// arrays on the device, the transposes of the boundary conditions, of
// solidCellBC and of the projection stage of poissonSolver, block
// snapshots for step checkpointing, and the C entry points of every
// tangent-linear and adjoint piece.
//
// A time step of a single block with a body (app/main.F90:133-263 in the
// order the drivers call it) is, at frozen geometry, x -> x' with
// x = (u, v, w, p, ut, vt, wt):
//   start (the geometry at the step's start, kinematics kin_a):
//     nsMomentum2order; velocityBC; solidCellBC; velocityForcing1(kin_a);
//     velocityBC; poissonSolver (projection with a fixed SOR count,
//     velocityBC, u <- ut); pressureForcing1(kin_a)
//   end (the geometry after the body moved, kinematics kin_b):
//     velocityForcingField(kin_b); pressureForcingField(kin_b)
// Only nsMomentum2order is non-linear in x; every other piece is linear in
// x and in the surface terms.
// ---------------------------------------------------------------------------
namespace bcfd {
namespace {

__global__ void k_ad_fill(double *__restrict__ x, int64_t n, double v) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) x[q] = v;
}

// transpose of velocityBC's z pass (k_bc_z), statements in reverse order
__global__ void k_adj_bc_z(Dims d, double *__restrict__ lu, double *__restrict__ lv, double *__restrict__ lw) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y;
  if (i > d.nx + 1) return;
  const int nz = d.nz;
  const int64_t c1 = ix3(d, i, j, 1), c2 = ix3(d, i, j, 2), cn1 = ix3(d, i, j, nz + 1), cn2 = ix3(d, i, j, nz + 2);
  lw[cn1] = 0.0;
  lv[cn1] += lv[cn2]; lv[cn2] = 0.0;
  lu[cn1] += lu[cn2]; lu[cn2] = 0.0;
  lw[c1] = 0.0;
  lv[c2] += lv[c1]; lv[c1] = 0.0;
  lu[c2] += lu[c1]; lu[c1] = 0.0;
}

// transpose of the y pass (k_bc_y)
__global__ void k_adj_bc_y(Dims d, double *__restrict__ lu, double *__restrict__ lv, double *__restrict__ lw) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int k = 2 + blockIdx.y;
  if (i > d.nx + 1) return;
  const int ny = d.ny;
  const int64_t c1 = ix3(d, i, 1, k), c2 = ix3(d, i, 2, k), cn1 = ix3(d, i, ny + 1, k), cn2 = ix3(d, i, ny + 2, k);
  lv[cn1] = 0.0;
  lw[cn1] += lw[cn2]; lw[cn2] = 0.0;
  lu[cn1] += lu[cn2]; lu[cn2] = 0.0;
  lv[c2] = 0.0;
  lw[c2] += lw[c1]; lw[c1] = 0.0;
  lu[c2] += lu[c1]; lu[c1] = 0.0;
}

// transpose of the x pass (k_bc_x): inflow, then the convective outflow,
// which reads the fields u, v, w of the step's start
__global__ void k_adj_bc_x(Dims d, double *__restrict__ lut, double *__restrict__ lvt, double *__restrict__ lwt,
                           double *__restrict__ lu, double *__restrict__ lv, double *__restrict__ lw,
                           const double *__restrict__ dxv, double dt, double uc) {
  const int j = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int k = 2 + blockIdx.y;
  if (j > d.ny + 1) return;
  const int nx = d.nx;
  const int64_t c1 = ix3(d, 1, j, k), cn = ix3(d, nx, j, k), cn1 = ix3(d, nx + 1, j, k), cn2 = ix3(d, nx + 2, j, k);
  const double c2 = (2.0 * dt / dxv[nx + 1]) * uc, cu = (dt / dxv[nx]) * uc;
  double a = lwt[cn2];
  lwt[cn2] = 0.0;
  lwt[cn1] -= a;
  lw[cn2] += a * (1.0 - c2);
  lw[cn1] += a * (1.0 + c2);
  a = lvt[cn2];
  lvt[cn2] = 0.0;
  lvt[cn1] -= a;
  lv[cn2] += a * (1.0 - c2);
  lv[cn1] += a * (1.0 + c2);
  a = lut[cn1];
  lut[cn1] = 0.0;
  lu[cn1] += a * (1.0 - cu);
  lu[cn] += a * cu;
  a = lwt[c1];
  lwt[c1] = 0.0;
  lwt[cn1] -= a;
  a = lvt[c1];
  lvt[c1] = 0.0;
  lvt[cn1] -= a;
  lut[c1] = 0.0;
}

// transpose of solidCellBC (k_solid_bc without the u copy): zero the
// adjoints of ut, vt, wt, p at solid cells
__global__ void k_adj_solid(Dims d, const uint8_t *__restrict__ cell, double *__restrict__ lut,
                            double *__restrict__ lvt, double *__restrict__ lwt, double *__restrict__ lp) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)d.nx * d.ny * d.nz) return;
  const int i = 2 + (int)(q % d.nx), j = 2 + (int)((q / d.nx) % d.ny), k = 2 + (int)(q / ((int64_t)d.nx * d.ny));
  const int64_t c = ix3(d, i, j, k);
  if (cell[c] != 1) return;
  lut[c] = 0.0; lvt[c] = 0.0; lwt[c] = 0.0; lp[c] = 0.0;
}

// transpose of u, v, w <- ut, vt, wt (every cell)
__global__ void k_adj_copy(int64_t n, double *__restrict__ lu, double *__restrict__ lv, double *__restrict__ lw,
                           double *__restrict__ lut, double *__restrict__ lvt, double *__restrict__ lwt) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  lut[q] += lu[q]; lu[q] = 0.0;
  lvt[q] += lv[q]; lv[q] = 0.0;
  lwt[q] += lw[q]; lw[q] = 0.0;
}

// snapshot of the state a time step reads and writes: fields, tags, the IB,
// TS and solid lists (in their capacity layouts), element and node data,
// the counts
struct Snapshot {
  bool live = false;
  std::vector<void *> bufs;
  std::vector<size_t> bytes;
  int64_t cap_ib = 0, cap_ts = 0, cap_solid = 0, n_mapped = 0;
  int64_t counts[BC_N] = {};
};
std::vector<Snapshot> g_snaps[kMaxBlocks + 1];

// The copies of one snapshot save or restore in one launch: grid row y
// copies bytes[y] from src[y] to dst[y] (device allocations, so both ends
// are 16-byte aligned), 16 bytes per thread and the tail byte by byte.
constexpr int kCopyMax = 64;
struct CopyList {
  const void *src[kCopyMax];
  void *dst[kCopyMax];
  int64_t bytes[kCopyMax];
};

__global__ void k_copy_list(CopyList L) {
  const int e = blockIdx.y;
  const char *__restrict__ s = (const char *)L.src[e];
  char *__restrict__ d = (char *)L.dst[e];
  const int64_t n = L.bytes[e], n16 = n / 16;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n16; q += stride)
    ((uint4 *)d)[q] = ((const uint4 *)s)[q];
  for (int64_t q = 16 * n16 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) d[q] = s[q];
}

// queue the copies, kCopyMax per launch
struct CopyQueue {
  CopyList L{};
  int n = 0;
  int64_t most = 0;
  void add(void *dst, const void *src, size_t bytes) {
    if (n == kCopyMax) flush();
    L.src[n] = src;
    L.dst[n] = dst;
    L.bytes[n] = (int64_t)bytes;
    most = (int64_t)bytes > most ? (int64_t)bytes : most;
    ++n;
  }
  int flush() {
    if (n == 0) return 0;
    const int64_t blocks = (most / 16 + 255) / 256;
    const unsigned gx = (unsigned)(blocks < 1 ? 1 : blocks > 512 ? 512 : blocks);
    hipLaunchKernelGGL(k_copy_list, dim3(gx, (unsigned)n), dim3(256), 0, rt().stream, L);
    n = 0;
    most = 0;
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
};

// every saved buffer of block b with its size at the current capacities
template <class F>
void snapshot_visit(Block &b, F f) {
  IbData &ib = b.ib;
  for (int q = 0; q < 7; ++q) f((void **)&b.f[q], sizeof(double) * (size_t)b.n2);
  for (int t = 0; t < BT_NTAG; ++t)
    f((void **)&b.tag[t], (size_t)((t == BT_CELL || t == BT_CELL_N) ? b.n2 : b.n3));
  if (!ib.present) return;
  f((void **)&ib.ib_map, sizeof(int) * (size_t)b.n2);
  if (ib.ib_ijk) {
    f((void **)&ib.ib_ijk, sizeof(int) * (size_t)(3 * ib.cap_ib));
    for (int q = 0; q < BN_NNEL; ++q) {
      f((void **)&ib.nel[q], sizeof(int) * (size_t)ib.cap_ib);
      f((void **)&ib.normdis[q], sizeof(double) * (size_t)ib.cap_ib);
    }
  }
  if (ib.ts_ijk) {
    f((void **)&ib.ts_ijk, sizeof(int) * (size_t)(3 * ib.cap_ts));
    f((void **)&ib.index_ts, sizeof(int) * (size_t)ib.cap_ts);
    for (int q = 0; q < BGH_N; ++q) f((void **)&ib.ghost[q], sizeof(double) * (size_t)ib.cap_ts);
  }
  if (ib.solid_ijk) f((void **)&ib.solid_ijk, sizeof(int) * (size_t)ib.cap_solid);
  for (int q = 0; q < 3; ++q) f((void **)&ib.node[q], sizeof(double) * (size_t)ib.n_nodes);
  for (int q = 0; q < BE_NELEM; ++q) f((void **)&ib.elem[q], sizeof(double) * (size_t)ib.n_elems);
}

int snapshot_save(Block &b, int id, int slot) {
  std::vector<Snapshot> &v = g_snaps[id];
  if ((int)v.size() <= slot) v.resize(slot + 1);
  Snapshot &sn = v[slot];
  IbData &ib = b.ib;
  size_t k = 0;
  int rc = 0;
  CopyQueue cq;
  snapshot_visit(b, [&](void **p, size_t n) {
    if (rc || !*p) return;
    if (k >= sn.bufs.size()) { sn.bufs.push_back(nullptr); sn.bytes.push_back(0); }
    if (sn.bytes[k] != n) {
      if (sn.bufs[k]) (void)hipFree(sn.bufs[k]);
      sn.bufs[k] = nullptr;
      char *q = nullptr;
      if (dev_alloc(&q, (int64_t)n)) { rc = -1; return; }
      sn.bufs[k] = q;
      sn.bytes[k] = n;
    }
    cq.add(sn.bufs[k], *p, n);
    ++k;
  });
  if (rc) return -1;
  if (cq.flush()) {
    set_error("snapshot: copy failed");
    return -1;
  }
  sn.cap_ib = ib.cap_ib; sn.cap_ts = ib.cap_ts; sn.cap_solid = ib.cap_solid; sn.n_mapped = ib.n_mapped;
  for (int q = 0; q < BC_N; ++q) sn.counts[q] = b.counts[q];
  sn.live = true;
  return 0;
}

int snapshot_restore(Block &b, int id, int slot) {
  std::vector<Snapshot> &v = g_snaps[id];
  if (slot < 0 || slot >= (int)v.size() || !v[slot].live) { set_error("snapshot %d of block %d not saved", slot, id); return -1; }
  Snapshot &sn = v[slot];
  IbData &ib = b.ib;
  // the list buffers at the saved capacities (their layouts depend on them)
  if (ib.present) {
    if (ib.cap_ib != sn.cap_ib) {
      if (dev_alloc(&ib.ib_ijk, 3 * sn.cap_ib)) return -1;
      for (int q = 0; q < BN_NNEL; ++q)
        if (dev_alloc(&ib.nel[q], sn.cap_ib) || dev_alloc(&ib.normdis[q], sn.cap_ib)) return -1;
      ib.cap_ib = sn.cap_ib;
    }
    if (ib.cap_ts != sn.cap_ts) {
      if (dev_alloc(&ib.ts_ijk, 3 * sn.cap_ts) || dev_alloc(&ib.index_ts, sn.cap_ts)) return -1;
      for (int q = 0; q < BGH_N; ++q)
        if (dev_alloc(&ib.ghost[q], sn.cap_ts)) return -1;
      ib.cap_ts = sn.cap_ts;
    }
    if (ib.cap_solid != sn.cap_solid) {
      if (dev_alloc(&ib.solid_ijk, sn.cap_solid)) return -1;
      ib.cap_solid = sn.cap_solid;
    }
  }
  size_t k = 0;
  int rc = 0;
  CopyQueue cq;
  snapshot_visit(b, [&](void **p, size_t n) {
    if (rc || !*p) return;
    if (k >= sn.bufs.size() || sn.bytes[k] != n) {
      set_error("snapshot %d of block %d does not match the block's buffers", slot, id);
      rc = -1;
      return;
    }
    cq.add(*p, sn.bufs[k], n);
    ++k;
  });
  if (rc) return -1;
  if (cq.flush()) {
    set_error("snapshot: copy failed");
    return -1;
  }
  ib.n_mapped = sn.n_mapped;
  for (int q = 0; q < BC_N; ++q) b.counts[q] = sn.counts[q];
  return 0;
}

void snapshot_free(int id, int slot) {
  std::vector<Snapshot> &v = g_snaps[id];
  for (int s = 0; s < (int)v.size(); ++s) {
    if (slot >= 0 && s != slot) continue;
    for (void *p : v[s].bufs)
      if (p) (void)hipFree(p);
    v[s] = Snapshot();
  }
  if (slot < 0) v.clear();
}

Block *ad_block(int id) {
  Block *b = get_block(id);
  if (!b) return nullptr;
  if (ad_ready(*b)) return nullptr;
  return b;
}

int n_interior_grid(const Block &b, dim3 &g) {
  g = dim3(div_up((int64_t)b.d.nx * b.d.ny * b.d.nz, 256));
  return 0;
}

int velocity_bc_adj(Block &b, double dt, double uc) {
  hipStream_t s = rt().stream;
  double **l = b.ad.s;
  hipLaunchKernelGGL(k_adj_bc_z, dim3(div_up(b.d.nx, 64), b.d.ny), dim3(64), 0, s, b.d, l[BF_UT], l[BF_VT], l[BF_WT]);
  hipLaunchKernelGGL(k_adj_bc_y, dim3(div_up(b.d.nx, 64), b.d.nz), dim3(64), 0, s, b.d, l[BF_UT], l[BF_VT], l[BF_WT]);
  hipLaunchKernelGGL(k_adj_bc_x, dim3(div_up(b.d.ny, 64), b.d.nz), dim3(64), 0, s, b.d, l[BF_UT], l[BF_VT], l[BF_WT],
                     l[BF_U], l[BF_V], l[BF_W], b.g[BG_DX], dt, uc);
  BCFD_CHECK_LAUNCH();
  return 0;
}

int velocity_bc_tlm(Block &b, double dt, double uc) {
  SwapFields sw(b);
  return velocity_bc(b, dt, uc, 0.0);
}

int solid_bc_adj(Block &b) {
  dim3 g;
  n_interior_grid(b, g);
  double **l = b.ad.s;
  hipLaunchKernelGGL(k_adj_solid, g, dim3(256), 0, rt().stream, b.d, b.tag[BT_CELL], l[BF_UT], l[BF_VT], l[BF_WT],
                     l[BF_P]);
  BCFD_CHECK_LAUNCH();
  return 0;
}

int solid_bc_tlm(Block &b) {
  // the forward zeroes ut, vt, wt, p at solid cells: so do the tangents
  return solid_bc_adj(b);
}

// tangent of poissonSolver on one block: the projection with n_iter SOR
// iterations, velocityBC (tangent inflow 0), u <- ut
int poisson_tlm(Block &b, int64_t n_iter, double dt, double uc) {
  SwapFields sw(b);
  if (projection(b, n_iter, dt)) return -1;
  if (velocity_bc(b, dt, uc, 0.0)) return -1;
  return norm_copy(b, dt);
}

int poisson_adj(Block &b, int64_t n_iter, double dt, double uc) {
  if (!b.A[0] || !b.g[BG_DX]) { set_error("poisson_adj: block is not set up"); return -1; }
  if (n_iter < 1) { set_error("poisson_adj: iteration count %lld < 1", (long long)n_iter); return -1; }
  hipStream_t s = rt().stream;
  double **l = b.ad.s;
  const int64_t n = b.n2;
  const dim3 blk(256), grd(div_up(n, 256));
  hipLaunchKernelGGL(k_adj_copy, grd, blk, 0, s, n, l[BF_U], l[BF_V], l[BF_W], l[BF_UT], l[BF_VT], l[BF_WT]);
  BCFD_CHECK_LAUNCH();
  if (velocity_bc_adj(b, dt, uc)) return -1;
  const Dims &d = b.d;
  const uint8_t *cell = b.tag[BT_CELL];
  double *lpc = b.ad.scr[0], *bbar = b.ad.scr[1];
  hipLaunchKernelGGL(k_adj_grad, grd, blk, 0, s, d, cell, l[BF_UT], l[BF_VT], l[BF_WT], l[BF_P], b.g[BG_DX],
                     b.g[BG_DY], b.g[BG_DZ], dt, lpc, bbar);
  if (adj_sor(b, n_iter, dt, b.ad.scr[2], lpc, bbar, &b.ad)) return -1;
  // each thread reads and writes only its own cell of the velocity adjoints
  hipLaunchKernelGGL(k_adj_div, grd, blk, 0, s, d, cell, bbar, l[BF_UT], l[BF_VT], l[BF_WT], b.g[BG_DX], b.g[BG_DY],
                     b.g[BG_DZ], l[BF_UT], l[BF_VT], l[BF_WT]);
  BCFD_CHECK_LAUNCH();
  return 0;
}

bool fp64_only() {
  if (precision_mode() != 0) { set_error("the tangent-linear and adjoint passes run the FP64 method only"); return false; }
  return true;
}

enum { AD_V1 = 0, AD_P1 = 1, AD_VF = 2, AD_PF = 3, AD_VG = 4, AD_PG = 5 };

// An element's centroid c = (p1 + p2 + p3) / 3, normal n = inor m / |m| and
// length |m|, m = (p2 - p1) x (p3 - p1) (computeSurfaceNorm,
// Search.F90:210-258): tangents from node tangents ...
__global__ void k_geom_tlm(int64_t ne, int64_t nn, const int *e1, const int *e2, const int *e3, const double *x,
                           const double *y, const double *z, const double *xd, const double *yd, const double *zd,
                           double inor, double *gdot) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int a = e1[e] - 1, b = e2[e] - 1, c = e3[e] - 1;
  const double p1[3] = {x[a], y[a], z[a]}, p2[3] = {x[b], y[b], z[b]}, p3[3] = {x[c], y[c], z[c]};
  const double d1[3] = {xd[a], yd[a], zd[a]}, d2[3] = {xd[b], yd[b], zd[b]}, d3[3] = {xd[c], yd[c], zd[c]};
  double u[3], v[3], du[3], dv[3], m[3], dm[3];
  for (int q = 0; q < 3; ++q) {
    u[q] = p2[q] - p1[q]; v[q] = p3[q] - p1[q];
    du[q] = d2[q] - d1[q]; dv[q] = d3[q] - d1[q];
    gdot[q * ne + e] = (d1[q] + d2[q] + d3[q]) / 3.0;
  }
  for (int q = 0; q < 3; ++q) {
    const int r = (q + 1) % 3, t = (q + 2) % 3;
    m[q] = u[r] * v[t] - u[t] * v[r];
    dm[q] = du[r] * v[t] - du[t] * v[r] + u[r] * dv[t] - u[t] * dv[r];
  }
  const double L = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  const double dL = (m[0] * dm[0] + m[1] * dm[1] + m[2] * dm[2]) / L;
  for (int q = 0; q < 3; ++q) gdot[(3 + q) * ne + e] = inor * (dm[q] / L - m[q] * dL / (L * L));
  gdot[6 * ne + e] = dL;
}

// ... and node adjoints from their adjoints (added atomically: nodes are shared)
__global__ void k_geom_adj(int64_t ne, int64_t nn, const int *e1, const int *e2, const int *e3, const double *x,
                           const double *y, const double *z, double inor, const double *gbar, double *nbar) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int a = e1[e] - 1, b = e2[e] - 1, c = e3[e] - 1;
  const double p1[3] = {x[a], y[a], z[a]}, p2[3] = {x[b], y[b], z[b]}, p3[3] = {x[c], y[c], z[c]};
  double u[3], v[3], m[3], cb[3], nb[3];
  for (int q = 0; q < 3; ++q) {
    u[q] = p2[q] - p1[q]; v[q] = p3[q] - p1[q];
    cb[q] = gbar[q * ne + e];
    nb[q] = gbar[(3 + q) * ne + e];
  }
  const double Lb = gbar[6 * ne + e];
  for (int q = 0; q < 3; ++q) {
    const int r = (q + 1) % 3, t = (q + 2) % 3;
    m[q] = u[r] * v[t] - u[t] * v[r];
  }
  const double L = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  const double mn = m[0] * nb[0] + m[1] * nb[1] + m[2] * nb[2];
  double mb[3], ub[3], vb[3];
  for (int q = 0; q < 3; ++q) mb[q] = inor * (nb[q] / L - m[q] * mn / (L * L * L)) + m[q] * Lb / L;
  // m = u x v: ubar = v x mbar, vbar = mbar x u
  for (int q = 0; q < 3; ++q) {
    const int r = (q + 1) % 3, t = (q + 2) % 3;
    ub[q] = v[r] * mb[t] - v[t] * mb[r];
    vb[q] = mb[r] * u[t] - mb[t] * u[r];
  }
  for (int q = 0; q < 3; ++q) {
    const double c3 = cb[q] / 3.0;
    atomicAdd(&nbar[q * nn + a], c3 - ub[q] - vb[q]);
    atomicAdd(&nbar[q * nn + b], c3 + ub[q]);
    atomicAdd(&nbar[q * nn + c], c3 + vb[q]);
  }
}

}  // namespace

int ad_geometry(Block &b) {
  AdWork &w = b.ad;
  if (w.gbar) return 0;
  if (!b.ib.present) { set_error("block has no immersed body"); return -1; }
  const int64_t ne = b.ib.n_elems, nn = b.ib.n_nodes;
  if (dev_alloc(&w.gbar, 7 * ne) || dev_alloc(&w.gdot, 7 * ne) || dev_alloc(&w.nbar, 3 * nn)) return -1;
  hipStream_t s = rt().stream;
  hipLaunchKernelGGL(k_ad_fill, dim3(div_up(7 * ne, 256)), dim3(256), 0, s, w.gbar, 7 * ne, 0.0);
  hipLaunchKernelGGL(k_ad_fill, dim3(div_up(7 * ne, 256)), dim3(256), 0, s, w.gdot, 7 * ne, 0.0);
  hipLaunchKernelGGL(k_ad_fill, dim3(div_up(3 * nn, 256)), dim3(256), 0, s, w.nbar, 3 * nn, 0.0);
  BCFD_CHECK_LAUNCH();
  return 0;
}

int ad_ready(Block &b) {
  AdWork &w = b.ad;
  if (w.s[0]) return 0;
  hipStream_t s = rt().stream;
  for (int q = 0; q < 7; ++q) {
    if (dev_alloc(&w.s[q], b.n2)) return -1;
    hipLaunchKernelGGL(k_ad_fill, dim3(div_up(b.n2, 256)), dim3(256), 0, s, w.s[q], b.n2, 0.0);
  }
  for (int q = 0; q < 3; ++q)
    if (dev_alloc(&w.scr[q], b.n2)) return -1;
  if (dev_alloc(&w.kin, 32)) return -1;
  hipLaunchKernelGGL(k_ad_fill, dim3(1), dim3(64), 0, s, w.kin, (int64_t)32, 0.0);
  BCFD_CHECK_LAUNCH();
  return 0;
}

void ad_release(Block &b) {
  AdWork &w = b.ad;
  if (w.sor_exec) (void)hipGraphExecDestroy(w.sor_exec);
  for (auto &p : w.s) if (p) { (void)hipFree(p); p = nullptr; }
  for (auto &p : w.scr) if (p) { (void)hipFree(p); p = nullptr; }
  for (auto &p : w.gh) if (p) { (void)hipFree(p); p = nullptr; }
  if (w.ybar) (void)hipFree(w.ybar);
  if (w.kin) (void)hipFree(w.kin);
  for (double *p : {w.addend, w.gbar, w.gdot, w.nbar})
    if (p) (void)hipFree(p);
  ad_release_records(b);
  w = AdWork();
  for (int id = 1; id <= kMaxBlocks; ++id)
    if (&rt().blk[id] == &b) snapshot_free(id, -1);
}

}  // namespace bcfd

extern "C" {

int biocfd_ad_reset(int id) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b) return -1;
  hipStream_t s = rt().stream;
  for (int q = 0; q < 7; ++q)
    hipLaunchKernelGGL(k_ad_fill, dim3(div_up(b->n2, 256)), dim3(256), 0, s, b->ad.s[q], b->n2, 0.0);
  for (int q = 0; q < BGH_N; ++q)
    if (b->ad.gh[q])
      hipLaunchKernelGGL(k_ad_fill, dim3(div_up(b->ad.cap_gh, 256)), dim3(256), 0, s, b->ad.gh[q], b->ad.cap_gh, 0.0);
  hipLaunchKernelGGL(k_ad_fill, dim3(1), dim3(64), 0, s, b->ad.kin, (int64_t)BK_N, 0.0);
  if (b->ad.gbar)
    hipLaunchKernelGGL(k_ad_fill, dim3(div_up(7 * b->ib.n_elems, 256)), dim3(256), 0, s, b->ad.gbar,
                       7 * b->ib.n_elems, 0.0);
  BCFD_CHECK_LAUNCH();
  return 0;
}

int biocfd_ad_upload(int id, int field, const double *host) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b) return -1;
  if (field < 0 || field > BF_P) { set_error("ad_upload: field %d outside u .. p", field); return -1; }
  return upload(b->ad.s[field], host, b->n2);
}

int biocfd_ad_download(int id, int field, double *host) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b) return -1;
  if (field < 0 || field > BF_P) { set_error("ad_download: field %d outside u .. p", field); return -1; }
  if (download(host, b->ad.s[field], b->n2)) return -1;
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  return 0;
}

int biocfd_ad_ghost_upload(int id, int which, const double *host, int64_t count) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b || ad_ghosts(*b)) return -1;
  if (which < 0 || which >= BGH_N || count < 0 || count > b->ad.cap_gh) {
    set_error("ad_ghost_upload: ghost %d, %lld entries (capacity %lld)", which, (long long)count,
              (long long)b->ad.cap_gh);
    return -1;
  }
  return count ? upload(b->ad.gh[which], host, count) : 0;
}

int biocfd_ad_ghost_download(int id, int which, double *host, int64_t count) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b || ad_ghosts(*b)) return -1;
  if (which < 0 || which >= BGH_N || count < 0 || count > b->ad.cap_gh) {
    set_error("ad_ghost_download: ghost %d, %lld entries (capacity %lld)", which, (long long)count,
              (long long)b->ad.cap_gh);
    return -1;
  }
  if (count && download(host, b->ad.gh[which], count)) return -1;
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  return 0;
}

int biocfd_ad_kin(int id, double kin_bar[BK_N], int clear) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b) return -1;
  if (download(rt().h_pinned, b->ad.kin, BK_N)) return -1;
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  for (int q = 0; q < BK_N; ++q) kin_bar[q] = rt().h_pinned[q];
  if (clear) {
    hipLaunchKernelGGL(k_ad_fill, dim3(1), dim3(64), 0, rt().stream, b->ad.kin, (int64_t)BK_N, 0.0);
    BCFD_CHECK_LAUNCH();
  }
  return 0;
}

int biocfd_tlm_momentum(int id, double dt, double re) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  return b ? momentum_tlm(*b, dt, re) : -1;
}

int biocfd_adj_momentum(int id, double dt, double re) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  return b ? momentum_adj(*b, dt, re) : -1;
}

int biocfd_tlm_velocity_bc(int id, double dt, double uc) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  return b ? velocity_bc_tlm(*b, dt, uc) : -1;
}

int biocfd_adj_velocity_bc(int id, double dt, double uc) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  return b ? velocity_bc_adj(*b, dt, uc) : -1;
}

int biocfd_tlm_solid_cell_bc(int id) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  return b ? solid_bc_tlm(*b) : -1;
}

int biocfd_adj_solid_cell_bc(int id) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  return b ? solid_bc_adj(*b) : -1;
}

int biocfd_tlm_forcing(int id, int routine, const double kin[BK_N], const double kin_dot[BK_N], int rec) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b) return -1;
  if (routine < AD_V1 || routine > AD_PG) { set_error("tlm_forcing: routine %d outside 0..5", routine); return -1; }
  return forcing_tlm(*b, routine, kin, kin_dot, rec);
}

int biocfd_adj_forcing(int id, int routine, const double kin[BK_N], int rec) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b) return -1;
  if (routine < AD_V1 || routine > AD_PG) { set_error("adj_forcing: routine %d outside 0..5", routine); return -1; }
  return forcing_adj(*b, routine, kin, rec);
}

int biocfd_tlm_poisson(int id, int64_t n_iter, double dt, double uc) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b || !fp64_only()) return -1;
  return poisson_tlm(*b, n_iter, dt, uc);
}

int biocfd_adj_poisson(int id, int64_t n_iter, double dt, double uc) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b || !fp64_only()) return -1;
  return poisson_adj(*b, n_iter, dt, uc);
}

int biocfd_stress_sums(int id, double mu_f, double rho_f, const double kin[BK_N], double out[4]) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  return b ? stress_sums(*b, mu_f, rho_f, kin, out) : -1;
}

int biocfd_tlm_stress(int id, double mu_f, double rho_f, const double kin[BK_N], const double kin_dot[BK_N],
                      int geom, double out[4]) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  return b ? stress_tlm(*b, mu_f, rho_f, kin, kin_dot, geom, out) : -1;
}

int biocfd_adj_stress(int id, double mu_f, double rho_f, const double kin[BK_N], const double sbar[4], int geom) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  return b ? stress_adj(*b, mu_f, rho_f, kin, sbar, geom) : -1;
}

int biocfd_tlm_step_start(int id, int64_t n_iter, double dt, double re, double uc, const double kin[BK_N],
                          const double kin_dot[BK_N], int rec) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b || !fp64_only()) return -1;
  if (momentum_tlm(*b, dt, re) || velocity_bc_tlm(*b, dt, uc) || solid_bc_tlm(*b) ||
      forcing_tlm(*b, AD_V1, kin, kin_dot, rec) || velocity_bc_tlm(*b, dt, uc) || poisson_tlm(*b, n_iter, dt, uc) ||
      forcing_tlm(*b, AD_P1, kin, kin_dot, rec))
    return -1;
  return 0;
}

int biocfd_tlm_step_end(int id, const double kin[BK_N], const double kin_dot[BK_N], int rec) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b) return -1;
  if (forcing_tlm(*b, AD_VF, kin, kin_dot, rec) || forcing_tlm(*b, AD_PF, kin, kin_dot, rec) ||
      forcing_tlm(*b, AD_VG, kin, kin_dot, rec) || forcing_tlm(*b, AD_PG, kin, kin_dot, rec))
    return -1;
  return 0;
}

int biocfd_adj_step_end(int id, const double kin[BK_N], int rec) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b) return -1;
  if (forcing_adj(*b, AD_PG, kin, rec) || forcing_adj(*b, AD_VG, kin, rec) || forcing_adj(*b, AD_PF, kin, rec) ||
      forcing_adj(*b, AD_VF, kin, rec))
    return -1;
  return 0;
}

int biocfd_adj_step_start(int id, int64_t n_iter, double dt, double re, double uc, const double kin[BK_N],
                          int rec) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b || !fp64_only()) return -1;
  if (forcing_adj(*b, AD_P1, kin, rec) || poisson_adj(*b, n_iter, dt, uc) || velocity_bc_adj(*b, dt, uc) ||
      forcing_adj(*b, AD_V1, kin, rec) || solid_bc_adj(*b) || velocity_bc_adj(*b, dt, uc) ||
      momentum_adj(*b, dt, re))
    return -1;
  return 0;
}

int biocfd_ad_record(int id, int slot) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b) return -1;
  if (slot < -1 || slot >= 4096) { set_error("ad_record: slot %d outside -1..4095", slot); return -1; }
  b->ad.rec_slot = slot;
  return 0;
}

int biocfd_ad_geometry_tangent(int id, double inor, const double *xdot, const double *ydot, const double *zdot) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b || ad_geometry(*b)) return -1;
  IbData &ib = b->ib;
  const int64_t ne = ib.n_elems, nn = ib.n_nodes;
  double *nd = b->ad.nbar;   // the node tangents travel through the node-adjoint buffer
  if (upload(nd, xdot, nn) || upload(nd + nn, ydot, nn) || upload(nd + 2 * nn, zdot, nn)) return -1;
  hipLaunchKernelGGL(k_geom_tlm, dim3(div_up(ne, 256)), dim3(256), 0, rt().stream, ne, nn, ib.el[0], ib.el[1],
                     ib.el[2], ib.node[0], ib.node[1], ib.node[2], nd, nd + nn, nd + 2 * nn, inor, b->ad.gdot);
  hipLaunchKernelGGL(k_ad_fill, dim3(div_up(3 * nn, 256)), dim3(256), 0, rt().stream, nd, 3 * nn, 0.0);
  BCFD_CHECK_LAUNCH();
  b->ad.gdot_set = true;
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  return 0;
}

int biocfd_ad_geometry_adjoint(int id, double inor, double *xbar, double *ybar, double *zbar, int clear) {
  BCFD_API_LOCK();
  Block *b = ad_block(id);
  if (!b || ad_geometry(*b)) return -1;
  IbData &ib = b->ib;
  const int64_t ne = ib.n_elems, nn = ib.n_nodes;
  double *nb = b->ad.nbar;
  hipStream_t s = rt().stream;
  hipLaunchKernelGGL(k_ad_fill, dim3(div_up(3 * nn, 256)), dim3(256), 0, s, nb, 3 * nn, 0.0);
  hipLaunchKernelGGL(k_geom_adj, dim3(div_up(ne, 256)), dim3(256), 0, s, ne, nn, ib.el[0], ib.el[1], ib.el[2],
                     ib.node[0], ib.node[1], ib.node[2], inor, b->ad.gbar, nb);
  BCFD_CHECK_LAUNCH();
  if (download(xbar, nb, nn) || download(ybar, nb + nn, nn) || download(zbar, nb + 2 * nn, nn)) return -1;
  if (clear) hipLaunchKernelGGL(k_ad_fill, dim3(div_up(7 * ne, 256)), dim3(256), 0, s, b->ad.gbar, 7 * ne, 0.0);
  BCFD_CHECK_LAUNCH();
  BCFD_HIP(hipStreamSynchronize(s));
  return 0;
}

int biocfd_snapshot_save(int id, int slot) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (slot < 0 || slot >= 4096) { set_error("snapshot slot %d outside 0..4095", slot); return -1; }
  return snapshot_save(*b, id, slot);
}

int biocfd_snapshot_restore(int id, int slot) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  return snapshot_restore(*b, id, slot);
}

int biocfd_snapshot_free(int id, int slot) {
  BCFD_API_LOCK();
  if (id < 1 || id > kMaxBlocks) { set_error("no block %d", id); return -1; }
  snapshot_free(id, slot);
  return 0;
}

}  // extern "C"
