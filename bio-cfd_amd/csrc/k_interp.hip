// Coarse/fine block coupling on gfx950: trilinear fine->coarse restriction
// (reference src/CoarseInterp.f90) and coarse->fine ghost-face prolongation
// (src/FineInterp_bound.f90), driven by the per-axis interface tables built
// on the host (src/InterfaceDetail.f90:168-241).  One thread per table entry
// (restriction) or per fine target cell (prolongation).
#include <cstring>

#include "biocfd_rt.hpp"

namespace bcfd {

// set_interface_detail: (coarse index, fine start, fine end) per coarse cell
// overlapped along one axis; first and last entries are the ghost entries
static void interface_axis(const std::vector<double> &cv, double lo, double hi, int64_t factor,
                           int64_t starter, std::vector<int64_t> &det, int64_t *counter) {
  const int64_t size = (int64_t)cv.size();
  int64_t first = 0, last = 0;
  for (int64_t i = 1; i <= size; ++i)
    if (cv[i - 1] > lo) { first = i; break; }
  for (int64_t i = first; i <= size; ++i)
    if (cv[i - 1] > hi) { last = i - 1; break; }
  det.clear();
  auto push = [&](int64_t a, int64_t b, int64_t c) { det.push_back(a); det.push_back(b); det.push_back(c); };
  push(first - 1, starter - 1, starter - 1);
  int64_t lo_f = starter, hi_f = starter + factor - 1;
  for (int64_t c = first;; ++c) {
    push(c, lo_f, hi_f);
    lo_f = hi_f + 1;
    hi_f = hi_f + factor;
    if (c + 1 > last) break;
  }
  push(last + 1, lo_f, lo_f);
  *counter = (int64_t)det.size() / 3;
}

struct ITab {
  const int *det[3];
  int cnt[3];
};

__device__ inline int DET(const ITab &t, int ax, int row, int n) { return t.det[ax][(row - 1) + 3 * (n - 1)]; }

__device__ inline double bl(double x, double y, double x1, double x2, double y1, double y2, double f1, double f2,
                            double f3, double f4) {
  const double numer = f1 * (x2 - x) * (y2 - y) + f3 * (x2 - x) * (y - y1) + f2 * (x - x1) * (y2 - y) +
                       f4 * (x - x1) * (y - y1);
  return numer / ((x2 - x1) * (y2 - y1));
}

struct GridSet {
  const double *x, *y, *z;
};

// fine -> coarse (CoarseInterp.f90:447-502 trilinear)
// entry q of the restriction of one (interface, component)
__device__ inline void restrict_at(const Dims &cd, const Dims &fd, const ITab &t, int comp, int cintp,
                                   const GridSet &cg, const GridSet &fg, const double *__restrict__ fvar,
                                   double *__restrict__ cvar, double *__restrict__ cvar2, int64_t q) {
  const int ni = t.cnt[0] - 2 * cintp + 1, nj = t.cnt[1] - 2 * cintp + 1, nk = t.cnt[2] - 2 * cintp + 1;
  if (ni <= 0 || nj <= 0 || nk <= 0) return;
  if (q >= (int64_t)ni * nj * nk) return;
  const int i = cintp + (int)(q % ni), j = cintp + (int)((q / ni) % nj), k = cintp + (int)(q / ((int64_t)ni * nj));
  const int tx = DET(t, 0, 1, i), ty = DET(t, 1, 1, j), tz = DET(t, 2, 1, k);
  int lx = (DET(t, 0, 2, i) + DET(t, 0, 3, i) - 1) / 2;
  int ly = (DET(t, 1, 2, j) + DET(t, 1, 3, j) - 1) / 2;
  int lz = (DET(t, 2, 2, k) + DET(t, 2, 3, k) - 1) / 2;
  int ox = 0, oy = 0, oz = 0;
  if (comp == 1) { lx = DET(t, 0, 3, i); ox = 1; }
  if (comp == 2) { ly = DET(t, 1, 3, j); oy = 1; }
  if (comp == 3) { lz = DET(t, 2, 3, k); oz = 1; }
  int ii = lx, jj = ly, kk = lz, vi = lx, vj = ly, vk = lz;
  if (comp == 1) { ii = lx - 1; vi = lx - 2; }
  else if (comp == 2) { jj = ly - 1; vj = ly - 2; }
  else if (comp == 3) { kk = lz - 1; vk = lz - 2; }
  if (vi < 1 || vj < 1 || vk < 1 || vi + 1 > fd.NX || vj + 1 > fd.NY || vk + 1 > fd.NZ || ii < 1 || jj < 1 ||
      kk < 1 || tx - ox < 1 || ty - oy < 1 || tz - oz < 1 || tx > cd.NX + 1 || ty > cd.NY + 1 || tz > cd.NZ + 1)
    return;  // interface outside the block: nothing to restrict here
  const double x = cg.x[tx - 1], y = cg.y[ty - 1], z = cg.z[tz - 1];
  const double X0 = fg.x[ii - 1], X1 = fg.x[ii], Y0 = fg.y[jj - 1], Y1 = fg.y[jj];
  const double z1 = bl(x, y, X0, X1, Y0, Y1, fvar[ix3(fd, vi, vj, vk)], fvar[ix3(fd, vi + 1, vj, vk)],
                       fvar[ix3(fd, vi, vj + 1, vk)], fvar[ix3(fd, vi + 1, vj + 1, vk)]);
  const double z2 = bl(x, y, X0, X1, Y0, Y1, fvar[ix3(fd, vi, vj, vk + 1)], fvar[ix3(fd, vi + 1, vj, vk + 1)],
                       fvar[ix3(fd, vi, vj + 1, vk + 1)], fvar[ix3(fd, vi + 1, vj + 1, vk + 1)]);
  const double Z0 = fg.z[kk - 1], Z1 = fg.z[kk];
  const double out = z1 + (z - Z0) * (z2 - z1) / (Z1 - Z0);
  const int64_t dst = ix3(cd, tx - ox, ty - oy, tz - oz);
  cvar[dst] = out;
  if (cvar2) cvar2[dst] = out;
}

__global__ void k_restrict(Dims cd, Dims fd, ITab t, int comp, int cintp, GridSet cg, GridSet fg,
                           const double *__restrict__ fvar, double *__restrict__ cvar, double *__restrict__ cvar2) {
  restrict_at(cd, fd, t, comp, cintp, cg, fg, fvar, cvar, cvar2, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// coarse -> fine on the six faces (FineInterp_bound.f90:554-603 trilinear
// over a two-coarse-cell span).  Entries: face-major enumeration of the
// (i,j,k) table entries with one index pinned to the first/last entry.
// element q of one axis pass of the prolongation of one (interface, component)
__device__ inline void prolong_at(const Dims &cd, const Dims &fd, const ITab &t, int comp, int axis,
                                  const GridSet &cg, const GridSet &fg, const double *__restrict__ cvar,
                                  double *__restrict__ fvar, double *__restrict__ fvar2, int factor, int64_t q) {
  // entries of this axis pass: the pinned axis takes {1, cnt}
  int n[3] = {t.cnt[0], t.cnt[1], t.cnt[2]};
  n[axis] = 2;
  const int64_t nent = (int64_t)n[0] * n[1] * n[2];
  const int per = factor * factor * factor;  // fine cells per entry (upper bound)
  if (q >= nent * per) return;
  const int64_t e = q / per;
  const int sub = (int)(q % per);
  int idx[3] = {(int)(e % n[0]), (int)((e / n[0]) % n[1]), (int)(e / ((int64_t)n[0] * n[1]))};
  for (int a = 0; a < 3; ++a) idx[a] += 1;
  if (idx[axis] == 2) idx[axis] = t.cnt[axis];
  const int i = idx[0], j = idx[1], k = idx[2];
  const int x1 = DET(t, 0, 2, i), x2 = DET(t, 0, 3, i);
  const int y1 = DET(t, 1, 2, j), y2 = DET(t, 1, 3, j);
  const int z1 = DET(t, 2, 2, k), z2 = DET(t, 2, 3, k);
  const int sx = sub % factor, sy = (sub / factor) % factor, sz = sub / (factor * factor);
  const int tx = x1 + sx, ty = y1 + sy, tz = z1 + sz;
  if (tx > x2 || ty > y2 || tz > z2) return;
  const int lx = DET(t, 0, 1, i), ly = DET(t, 1, 1, j), lz = DET(t, 2, 1, k);
  int vi = lx, vj = ly, vk = lz;
  if (comp == 1) vi = lx - 1;
  else if (comp == 2) vj = ly - 1;
  else if (comp == 3) vk = lz - 1;
  const int ox_ = (comp == 1), oy_ = (comp == 2), oz_ = (comp == 3);
  if (vi - 1 < 1 || vj - 1 < 1 || vk - 1 < 1 || vi + 1 > cd.NX || vj + 1 > cd.NY || vk + 1 > cd.NZ || lx < 2 ||
      ly < 2 || lz < 2 || tx - ox_ < 1 || ty - oy_ < 1 || tz - oz_ < 1 || tx - ox_ > fd.NX || ty - oy_ > fd.NY ||
      tz - oz_ > fd.NZ)
    return;  // the fine block touches the coarse block's edge: no coarse stencil
  const double x = fg.x[tx - 1], y = fg.y[ty - 1], z = fg.z[tz - 1];
  const double X0 = cg.x[lx - 2], X1 = cg.x[lx], Y0 = cg.y[ly - 2], Y1 = cg.y[ly];
  const double za = bl(x, y, X0, X1, Y0, Y1, cvar[ix3(cd, vi - 1, vj - 1, vk - 1)], cvar[ix3(cd, vi + 1, vj - 1, vk - 1)],
                       cvar[ix3(cd, vi - 1, vj + 1, vk - 1)], cvar[ix3(cd, vi + 1, vj + 1, vk - 1)]);
  const double zb = bl(x, y, X0, X1, Y0, Y1, cvar[ix3(cd, vi - 1, vj - 1, vk + 1)], cvar[ix3(cd, vi + 1, vj - 1, vk + 1)],
                       cvar[ix3(cd, vi - 1, vj + 1, vk + 1)], cvar[ix3(cd, vi + 1, vj + 1, vk + 1)]);
  const double Z0 = cg.z[lz - 2], Z1 = cg.z[lz];
  const double out = za + (z - Z0) * (zb - za) / (Z1 - Z0);
  const int ox = (comp == 1), oy = (comp == 2), oz = (comp == 3);
  const int64_t dst = ix3(fd, tx - ox, ty - oy, tz - oz);
  fvar[dst] = out;
  if (fvar2) fvar2[dst] = out;
}

__global__ void k_prolong(Dims cd, Dims fd, ITab t, int comp, int axis, GridSet cg, GridSet fg,
                          const double *__restrict__ cvar, double *__restrict__ fvar, double *__restrict__ fvar2,
                          int factor) {
  prolong_at(cd, fd, t, comp, axis, cg, fg, cvar, fvar, fvar2, factor,
             (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// One launch for the same update of every live interface (and component):
// task b covers threads first[b] .. first[b + 1] - 1 (first[ntask] = the
// launch's threads).  For the prolongation, one launch per axis pass.
struct InterpTask {
  Dims cd, fd;
  ITab t;
  int comp, cintp, factor;
  GridSet cg, fg;
  const double *src;
  double *dst;
  int64_t first[3];   // per axis pass (the restriction uses first[0])
};

__device__ inline int task_of(const InterpTask *__restrict__ tasks, int ntask, int axis, int64_t q) {
  int lo = 0, hi = ntask - 1;   // the last task whose first <= q
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (tasks[mid].first[axis] <= q) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void k_prolong_set(const InterpTask *__restrict__ tasks, int ntask, int axis, int64_t total) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const InterpTask &T = tasks[task_of(tasks, ntask, axis, q)];
  prolong_at(T.cd, T.fd, T.t, T.comp, axis, T.cg, T.fg, T.src, T.dst, nullptr, T.factor, q - T.first[axis]);
}

__global__ void k_restrict_set(const InterpTask *__restrict__ tasks, int ntask, int64_t total) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const InterpTask &T = tasks[task_of(tasks, ntask, 0, q)];
  restrict_at(T.cd, T.fd, T.t, T.comp, T.cintp, T.cg, T.fg, T.src, T.dst, nullptr, q - T.first[0]);
}

// coarse -> fine volume fill after a block move (FineInterp.f90:10-281):
// each fine cell of a table entry (entries 2 .. counter-1 per axis) whose
// cell_n is not 1 takes the bilinear value, in x and y, of the coarse planes
// below and above the entry's coarse cell, blended linearly in z.  The
// staggered component's coarse values and its fine target sit one index
// lower along its axis (u(i) at xu(i+1)).  One thread per (entry, fine cell).
__global__ void k_fine_mv(Dims cd, Dims fd, ITab t, int comp, GridSet cg, GridSet fg,
                          const uint8_t *__restrict__ cell_n, const double *__restrict__ cvar,
                          double *__restrict__ fvar, int factor) {
  const int ni = t.cnt[0] - 2, nj = t.cnt[1] - 2, nk = t.cnt[2] - 2;
  if (ni <= 0 || nj <= 0 || nk <= 0) return;
  const int per = factor * factor * factor;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)ni * nj * nk * per) return;
  const int64_t e = q / per;
  const int sub = (int)(q % per);
  const int i = 2 + (int)(e % ni), j = 2 + (int)((e / ni) % nj), k = 2 + (int)(e / ((int64_t)ni * nj));
  const int s = DET(t, 0, 2, i) + sub % factor, l = DET(t, 1, 2, j) + (sub / factor) % factor,
            r = DET(t, 2, 2, k) + sub / (factor * factor);
  if (s > DET(t, 0, 3, i) || l > DET(t, 1, 3, j) || r > DET(t, 2, 3, k)) return;
  if (cell_n[ix3(fd, s, l, r)] == 1) return;
  const int lx = DET(t, 0, 1, i), ly = DET(t, 1, 1, j), lz = DET(t, 2, 1, k);
  const int ox = comp == 1, oy = comp == 2, oz = comp == 3;
  const int i0 = lx - 1 - ox, i1 = lx + 1 - ox, j0 = ly - 1 - oy, j1 = ly + 1 - oy, k0 = lz - 1 - oz,
            k1 = lz + 1 - oz;
  const double vx = fg.x[s - 1], vy = fg.y[l - 1], vz = fg.z[r - 1];
  const double x1 = cg.x[lx - 2], x2 = cg.x[lx], y1 = cg.y[ly - 2], y2 = cg.y[ly], z1 = cg.z[lz - 2],
               z2 = cg.z[lz];
  const double deno = (x2 - x1) * (y2 - y1);
  const double xtx = (x2 - vx), xxo = (vx - x1), yty = (y2 - vy), yyo = (vy - y1);
  double first = (cvar[ix3(cd, i0, j0, k0)] * xtx + cvar[ix3(cd, i1, j0, k0)] * xxo) * yty;
  double second = (cvar[ix3(cd, i0, j1, k0)] * xtx + cvar[ix3(cd, i1, j1, k0)] * xxo) * yyo;
  const double ans1 = (first + second) / deno;
  first = (cvar[ix3(cd, i0, j0, k1)] * xtx + cvar[ix3(cd, i1, j0, k1)] * xxo) * yty;
  second = (cvar[ix3(cd, i0, j1, k1)] * xtx + cvar[ix3(cd, i1, j1, k1)] * xxo) * yyo;
  const double ans2 = (first + second) / deno;
  fvar[ix3(fd, s - ox, l - oy, r - oz)] = ans1 + ((vz - z1) * ((ans2 - ans1) / (z2 - z1)));
}

static GridSet grids_of(const Block &b, int comp) {
  switch (comp) {
    case 0: return {b.g[BG_XP], b.g[BG_YP], b.g[BG_ZP]};
    case 1: return {b.g[BG_XU], b.g[BG_YU], b.g[BG_ZU]};
    case 2: return {b.g[BG_XV], b.g[BG_YV], b.g[BG_ZV]};
    default: return {b.g[BG_XW], b.g[BG_YW], b.g[BG_ZW]};
  }
}

static ITab tab_of(const Interface &it, int comp) {
  ITab t;
  for (int a = 0; a < 3; ++a) { t.det[a] = it.det[comp][a]; t.cnt[a] = (int)it.counter[comp][a]; }
  return t;
}

// The fields an update of kind `which` moves, per component: 0 = p,u,v,w ;
// 1 = ut,vt,wt ; 2 = pc
static int update_fields(int which, int comp) {
  static const int f0[4] = {BF_P, BF_U, BF_V, BF_W}, f1[4] = {-1, BF_UT, BF_VT, BF_WT};
  if (which == 0) return f0[comp];
  if (which == 1) return f1[comp];
  return comp == 0 ? BF_PC : -1;
}

// Whether every live interface's update can run in one launch: each fine
// block is the fine side of one interface only and no block is both a fine
// and a coarse side (so no update reads what another writes); for the
// restriction also the coarse cells the interfaces write (per component)
// must not overlap.
static bool batchable(bool restrict_kind) {
  int as_b[kMaxBlocks + 1] = {}, as_a[kMaxBlocks + 1] = {};
  for (int q = 1; q <= kMaxInterfaces; ++q) {
    const Interface &it = rt().itf[q];
    if (!it.live) continue;
    if (++as_b[it.b] > 1) return false;
    ++as_a[it.a];
  }
  for (int g = 1; g <= kMaxBlocks; ++g)
    if (as_a[g] && as_b[g]) return false;
  if (!restrict_kind) return true;
  // coarse index boxes [lo, hi] per axis of the cells each restriction writes
  struct Box { int a, comp; int64_t lo[3], hi[3]; };
  std::vector<Box> boxes;
  for (int q = 1; q <= kMaxInterfaces; ++q) {
    const Interface &it = rt().itf[q];
    if (!it.live) continue;
    const int64_t cintp = rt().blk[it.b].cintp;
    for (int comp = 0; comp < 4; ++comp) {
      Box bx{it.a, comp, {0, 0, 0}, {0, 0, 0}};
      bool empty = false;
      for (int ax = 0; ax < 3; ++ax) {
        const int64_t cnt = it.counter[comp][ax];
        const std::vector<int64_t> &h = it.hdet[comp][ax];
        const int64_t n0 = cintp, n1 = cnt - cintp;
        if (n1 < n0 || n0 < 1 || (size_t)(3 * n1) > h.size()) { empty = true; break; }
        const int o = comp == ax + 1;
        bx.lo[ax] = h[3 * (n0 - 1)] - o;
        bx.hi[ax] = h[3 * (n1 - 1)] - o;
      }
      if (empty) continue;
      for (const Box &o : boxes) {
        if (o.a != bx.a || o.comp != bx.comp) continue;
        bool apart = false;
        for (int ax = 0; ax < 3; ++ax) apart = apart || bx.hi[ax] < o.lo[ax] || o.hi[ax] < bx.lo[ax];
        if (!apart) return false;
      }
      boxes.push_back(bx);
    }
  }
  return true;
}

// device task tables of the batched updates, per kind (0 restriction, 1
// prolongation) and `which`, re-uploaded when their contents change
struct TaskTable {
  std::vector<InterpTask> h;
  InterpTask *d = nullptr;
  int64_t cap = 0;
  int64_t total[3] = {0, 0, 0};
};
static TaskTable g_tasks[2][3];

static int upload_tasks(TaskTable &tt, std::vector<InterpTask> &v) {
  if (v.size() == tt.h.size() && (v.empty() || !memcmp(v.data(), tt.h.data(), sizeof(InterpTask) * v.size())))
    return 0;
  hipStream_t s = rt().stream;
  // a launch queued earlier may still read the old table
  BCFD_HIP(hipStreamSynchronize(s));
  if (grow(&tt.d, &tt.cap, (int64_t)v.size())) return -1;
  tt.h = v;
  if (!v.empty())
    BCFD_HIP(hipMemcpyAsync(tt.d, tt.h.data(), sizeof(InterpTask) * v.size(), hipMemcpyHostToDevice, s));
  BCFD_HIP(hipStreamSynchronize(s));
  return 0;
}

// the updates of kind `which` of every live interface in one launch per
// pass (see batchable)
static int batched_update(bool restrict_kind, int which) {
  std::vector<InterpTask> v;
  int64_t total[3] = {0, 0, 0};
  for (int q = 1; q <= kMaxInterfaces; ++q) {
    Interface &it = rt().itf[q];
    if (!it.live) continue;
    Block &a = rt().blk[it.a], &b = rt().blk[it.b];
    for (int comp = 0; comp < 4; ++comp) {
      const int f = update_fields(which, comp);
      if (f < 0) continue;
      InterpTask T;
      memset(&T, 0, sizeof T);
      T.cd = a.d;
      T.fd = b.d;
      T.t = tab_of(it, comp);
      T.comp = comp;
      T.cintp = (int)b.cintp;
      T.factor = (int)(it.b_msh / it.a_msh);
      T.cg = grids_of(a, comp);
      T.fg = grids_of(b, comp);
      T.src = restrict_kind ? b.f[f] : a.f[f];
      T.dst = restrict_kind ? a.f[f] : b.f[f];
      if (restrict_kind) {
        const int64_t n = (int64_t)(T.t.cnt[0] - 2 * T.cintp + 1) * (T.t.cnt[1] - 2 * T.cintp + 1) *
                          (T.t.cnt[2] - 2 * T.cintp + 1);
        if (n <= 0) continue;
        T.first[0] = total[0];
        total[0] += n;
      } else {
        for (int axis = 0; axis < 3; ++axis) {
          int n[3] = {T.t.cnt[0], T.t.cnt[1], T.t.cnt[2]};
          n[axis] = 2;
          T.first[axis] = total[axis];
          total[axis] += (int64_t)n[0] * n[1] * n[2] * T.factor * T.factor * T.factor;
        }
      }
      v.push_back(T);
    }
  }
  if (v.empty()) return 0;
  TaskTable &tt = g_tasks[restrict_kind ? 0 : 1][which];
  if (upload_tasks(tt, v)) return -1;
  hipStream_t s = rt().stream;
  const int nt = (int)v.size();
  if (restrict_kind) {
    hipLaunchKernelGGL(k_restrict_set, dim3(div_up(total[0], 256)), dim3(256), 0, s, tt.d, nt, total[0]);
  } else {
    for (int axis = 0; axis < 3; ++axis)
      if (total[axis] > 0)
        hipLaunchKernelGGL(k_prolong_set, dim3(div_up(total[axis], 256)), dim3(256), 0, s, tt.d, nt, axis,
                           total[axis]);
  }
  BCFD_CHECK_LAUNCH();
  return 0;
}

// one launch per pass for all interfaces where allowed (the default), else
// interface by interface
static int g_interp_batch = 1;

int interface_coarse_update(int which) {
  if (g_interp_batch && batchable(true)) return batched_update(true, which);
  hipStream_t s = rt().stream;
  for (int q = 1; q <= kMaxInterfaces; ++q) {
    Interface &it = rt().itf[q];
    if (!it.live) continue;
    Block &a = rt().blk[it.a], &b = rt().blk[it.b];
    const int cintp = (int)b.cintp;
    auto run = [&](int comp, const double *fv, double *cv, double *cv2) {
      ITab t = tab_of(it, comp);
      const int64_t n = (int64_t)(t.cnt[0] - 2 * cintp + 1) * (t.cnt[1] - 2 * cintp + 1) * (t.cnt[2] - 2 * cintp + 1);
      if (n <= 0) return;
      hipLaunchKernelGGL(k_restrict, dim3(div_up(n, 256)), dim3(256), 0, s, a.d, b.d, t, comp, cintp,
                         grids_of(a, comp), grids_of(b, comp), fv, cv, cv2);
    };
    if (which == 0) {
      run(0, b.f[BF_P], a.f[BF_P], nullptr);
      run(1, b.f[BF_U], a.f[BF_U], nullptr);
      run(2, b.f[BF_V], a.f[BF_V], nullptr);
      run(3, b.f[BF_W], a.f[BF_W], nullptr);
    } else if (which == 1) {
      run(1, b.f[BF_UT], a.f[BF_UT], nullptr);
      run(2, b.f[BF_VT], a.f[BF_VT], nullptr);
      run(3, b.f[BF_WT], a.f[BF_WT], nullptr);
    } else {
      run(0, b.f[BF_PC], a.f[BF_PC], nullptr);
    }
  }
  BCFD_CHECK_LAUNCH();
  return 0;
}

int interface_fine_update_one(int q, int which) {
  hipStream_t s = rt().stream;
  {
    Interface &it = rt().itf[q];
    Block &a = rt().blk[it.a], &b = rt().blk[it.b];
    const int factor = (int)(it.b_msh / it.a_msh);
    auto run = [&](int comp, const double *cv, double *fv) {
      ITab t = tab_of(it, comp);
      for (int axis = 0; axis < 3; ++axis) {
        int n[3] = {t.cnt[0], t.cnt[1], t.cnt[2]};
        n[axis] = 2;
        const int64_t total = (int64_t)n[0] * n[1] * n[2] * factor * factor * factor;
        hipLaunchKernelGGL(k_prolong, dim3(div_up(total, 256)), dim3(256), 0, s, a.d, b.d, t, comp, axis,
                           grids_of(a, comp), grids_of(b, comp), cv, fv, (double *)nullptr, factor);
      }
    };
    if (which == 0) {
      run(0, a.f[BF_P], b.f[BF_P]);
      run(1, a.f[BF_U], b.f[BF_U]);
      run(2, a.f[BF_V], b.f[BF_V]);
      run(3, a.f[BF_W], b.f[BF_W]);
    } else if (which == 1) {
      run(1, a.f[BF_UT], b.f[BF_UT]);
      run(2, a.f[BF_VT], b.f[BF_VT]);
      run(3, a.f[BF_WT], b.f[BF_WT]);
    } else {
      run(0, a.f[BF_PC], b.f[BF_PC]);
    }
  }
  BCFD_CHECK_LAUNCH();
  return 0;
}

int interface_fine_update(int which) {
  if (g_interp_batch && batchable(false)) return batched_update(false, which);
  for (int q = 1; q <= kMaxInterfaces; ++q)
    if (rt().itf[q].live && interface_fine_update_one(q, which)) return -1;
  return 0;
}

// fineUpdate_mv of interface q.  The reference reads the coarse planes and
// rows at loc-1 .. loc+1 (loc-2 for the staggered component) without a
// check; here every table entry is checked on the host first, and an entry
// whose stencil leaves either block is an error.
int interface_fine_update_mv(int q) {
  Interface &it = rt().itf[q];
  Block &a = rt().blk[it.a], &b = rt().blk[it.b];
  const int factor = (int)(it.b_msh / it.a_msh);
  const int cn[3] = {a.d.NX, a.d.NY, a.d.NZ}, fn[3] = {b.d.NX, b.d.NY, b.d.NZ};
  for (int comp = 0; comp < 4; ++comp)
    for (int ax = 0; ax < 3; ++ax) {
      const std::vector<int64_t> &h = it.hdet[comp][ax];
      const int64_t cnt = it.counter[comp][ax];
      const int off = comp == ax + 1;
      for (int64_t n = 2; n <= cnt - 1; ++n) {
        const int64_t loc = h[3 * (n - 1)], lo = h[3 * (n - 1) + 1], hi = h[3 * (n - 1) + 2];
        if (loc - 1 - off < 1 || loc + 1 > cn[ax] || lo - off < 1 || hi > fn[ax]) {
          set_error("fine_update_mv: interface %d entry %lld reaches outside the blocks", q, (long long)n);
          return -1;
        }
      }
    }
  hipStream_t s = rt().stream;
  auto run = [&](int comp, const double *cv, double *fv) {
    ITab t = tab_of(it, comp);
    const int64_t total = (int64_t)(t.cnt[0] - 2) * (t.cnt[1] - 2) * (t.cnt[2] - 2) * factor * factor * factor;
    if (t.cnt[0] <= 2 || t.cnt[1] <= 2 || t.cnt[2] <= 2) return;
    hipLaunchKernelGGL(k_fine_mv, dim3(div_up(total, 256)), dim3(256), 0, s, a.d, b.d, t, comp, grids_of(a, comp),
                       grids_of(b, comp), b.tag[BT_CELL_N], cv, fv, factor);
  };
  run(0, a.f[BF_P], b.f[BF_P]);
  run(1, a.f[BF_U], b.f[BF_U]);
  run(2, a.f[BF_V], b.f[BF_V]);
  run(3, a.f[BF_W], b.f[BF_W]);
  BCFD_CHECK_LAUNCH();
  return 0;
}

}  // namespace bcfd

using namespace bcfd;

extern "C" {

int biocfd_interp_batching(int on) {
  BCFD_API_LOCK();
  g_interp_batch = on ? 1 : 0;
  return 0;
}

}  // extern "C"

extern "C" {

int biocfd_interface_create(int itf, int a_blk, int b_blk, int64_t a_msh, int64_t b_msh, const double ext[6]) {
  BCFD_API_LOCK();
  if (ensure_init()) return -1;
  if (itf < 1 || itf > kMaxInterfaces) { set_error("interface id %d out of range", itf); return -1; }
  Block *a = get_block(a_blk);
  Block *b = get_block(b_blk);
  if (!a || !b) return -1;
  if (a_msh <= 0 || b_msh % a_msh != 0) { set_error("interface: refinement ratio must be an integer"); return -1; }
  Interface &it = rt().itf[itf];
  for (auto &row : it.det)
    for (auto &p : row)
      if (p) { (void)hipFree(p); p = nullptr; }
  it = Interface();
  it.a = a_blk; it.b = b_blk; it.a_msh = a_msh; it.b_msh = b_msh;
  for (int q = 0; q < 6; ++q) it.ext[q] = ext[q];
  const int64_t factor = b_msh / a_msh;
  const int gx[4] = {BG_XP, BG_XU, BG_XV, BG_XW}, gy[4] = {BG_YP, BG_YU, BG_YV, BG_YW},
            gz[4] = {BG_ZP, BG_ZU, BG_ZV, BG_ZW};
  for (int c = 0; c < 4; ++c) {
    const int grid[3] = {gx[c], gy[c], gz[c]};
    for (int ax = 0; ax < 3; ++ax) {
      if (a->hg[grid[ax]].empty()) { set_error("interface: coarse grids not uploaded"); return -1; }
      const int64_t starter = (c == ax + 1) ? 3 : 2;  // normal velocity component
      interface_axis(a->hg[grid[ax]], ext[2 * ax], ext[2 * ax + 1], factor, starter, it.hdet[c][ax],
                     &it.counter[c][ax]);
      std::vector<int> t32(it.hdet[c][ax].begin(), it.hdet[c][ax].end());
      if (dev_alloc(&it.det[c][ax], (int64_t)t32.size())) return -1;
      BCFD_HIP(hipMemcpy(it.det[c][ax], t32.data(), sizeof(int) * t32.size(), hipMemcpyHostToDevice));
    }
  }
  it.live = true;
  return 0;
}

int biocfd_interface_table(int itf, int comp, int axis, int64_t *host, int64_t *counter, int64_t cap) {
  BCFD_API_LOCK();
  if (ensure_init()) return -1;
  if (itf < 1 || itf > kMaxInterfaces || !rt().itf[itf].live || comp < 0 || comp > 3 || axis < 0 || axis > 2) {
    set_error("interface_table: bad request");
    return -1;
  }
  const Interface &it = rt().itf[itf];
  *counter = it.counter[comp][axis];
  const std::vector<int64_t> &v = it.hdet[comp][axis];
  if ((int64_t)v.size() > 3 * cap) { set_error("interface_table: capacity too small"); return -1; }
  for (size_t q = 0; q < v.size(); ++q) host[q] = v[q];
  return 0;
}

int biocfd_coarse_update(int which) {
  BCFD_API_LOCK();
  if (ensure_init()) return -1;
  if (which < 0 || which > 2) { set_error("coarse_update: bad selector %d", which); return -1; }
  return interface_coarse_update(which);
}

int biocfd_fine_update_mv(int itf) {
  BCFD_API_LOCK();
  if (ensure_init()) return -1;
  if (itf < 1 || itf > kMaxInterfaces || !rt().itf[itf].live) {
    set_error("fine_update_mv: no interface %d", itf);
    return -1;
  }
  return interface_fine_update_mv(itf);
}

int biocfd_fine_update_bd(int which) {
  BCFD_API_LOCK();
  if (ensure_init()) return -1;
  if (which < 0 || which > 2) { set_error("fine_update_bd: bad selector %d", which); return -1; }
  return interface_fine_update(which);
}

int biocfd_fine_update_bd_one(int itf, int which) {
  BCFD_API_LOCK();
  if (ensure_init()) return -1;
  if (which < 0 || which > 2) { set_error("fine_update_bd_one: bad selector %d", which); return -1; }
  if (itf < 1 || itf > kMaxInterfaces || !rt().itf[itf].live) {
    set_error("fine_update_bd_one: no interface %d", itf);
    return -1;
  }
  return interface_fine_update_one(itf, which);
}

}  // extern "C"
