// Block move on gfx950: the reference moves a fine block by whole coarse
// cells when its body nears the block's faces (src/Search.F90:
// findDistnode :21-47, cellCount_solid_coarse_mv :916-995, block_move_check
// :997-1203, change_block_coords :1208-1264, change_block_coords_interfaces
// :1269-1370, change_block_interface :1372-1461, tagging_th_move :415-432;
// src/FineInterp.f90 fineUpdate_mv :10-281 in k_interp.hip).
//
// The decisions, the new extents and the grid shift are host arithmetic on
// the host copies of the grids, in the reference's expression order; the
// field work runs on the device: keeping u, v, w, p, the interior mask
// cell_n of the moved block and the shifted copy of the kept overlap (one
// pass that also zeroes the rest of the block, as the reference zeroes the
// fields before copying).
#include <cmath>

#include "biocfd_rt.hpp"

namespace bcfd {

int interface_fine_update_one(int itf, int which);
int interface_fine_update_mv(int itf);

namespace {

// kinematic state slots (include/biocfd_hip.h, BKS_N)
constexpr int S_PIVZ = 25, S_INITY = 29, S_NXTY = 30, S_INITX = 31, S_NXTX = 32, S_INITZ = 33, S_NXTZ = 34;

// a shift of m cells made a whole number of coarse cells, at least one
// (Search.F90:1043-1059)
int64_t whole_coarse(int64_t m, int64_t factor) {
  const int64_t am = m < 0 ? -m : m;
  if (am < factor) return m < 0 ? -factor : factor;
  return m < 0 ? m + am % factor : m - am % factor;
}

// first centre above lo, then the last centre not above hi after it; a bound
// no centre meets keeps its previous value (Search.F90:1129-1184)
void copy_range(const std::vector<double> &c, double lo, double hi, int64_t *start, int64_t *end) {
  const int64_t n = (int64_t)c.size();
  for (int64_t i = 1; i <= n; ++i)
    if (c[i - 1] > lo) { *start = i; break; }
  for (int64_t i = *start < 1 ? 1 : *start; i <= n; ++i)
    if (c[i - 1] > hi) { *end = i - 1; break; }
}

double dmax(double a, double b) { return a > b ? a : b; }
double dmin(double a, double b) { return a < b ? a : b; }

// cell_n = 1 strictly inside the kept overlap shrunk by one cell, over the
// whole (n+2)^3 block (Search.F90:1281-1294)
__global__ void k_interior_mask(Dims d, const double *__restrict__ xp, const double *__restrict__ yp,
                                const double *__restrict__ zp, double xs, double xe, double ys, double ye, double zs,
                                double ze, uint8_t *__restrict__ cell_n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > d.NX) return;
  const double x = xp[i - 1], y = yp[j - 1], z = zp[k - 1];
  cell_n[ix3(d, i, j, k)] = (x > xs && x < xe && z > zs && z < ze && y > ys && y < ye) ? 1 : 0;
}

// the kept overlap copied to its new indices, zero elsewhere
// (Search.F90:1345-1368): target (i,j,k) in [lo, hi] reads (i,j,k) - lo + src
struct Box {
  int lo[3], hi[3], src[3];
};

__global__ void k_shift_fields(Dims d, Box bx, const double *__restrict__ du, const double *__restrict__ dv,
                               const double *__restrict__ dw, const double *__restrict__ dp, double *__restrict__ u,
                               double *__restrict__ v, double *__restrict__ w, double *__restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > d.NX) return;
  const int64_t t = ix3(d, i, j, k);
  if (i >= bx.lo[0] && i <= bx.hi[0] && j >= bx.lo[1] && j <= bx.hi[1] && k >= bx.lo[2] && k <= bx.hi[2]) {
    const int64_t s = ix3(d, i - bx.lo[0] + bx.src[0], j - bx.lo[1] + bx.src[1], k - bx.lo[2] + bx.src[2]);
    u[t] = du[s];
    v[t] = dv[s];
    w[t] = dw[s];
    p[t] = dp[s];
  } else {
    u[t] = 0.0;
    v[t] = 0.0;
    w[t] = 0.0;
    p[t] = 0.0;
  }
}

Interface *live_interface(int itf, const char *who) {
  if (ensure_init()) return nullptr;
  if (itf < 1 || itf > kMaxInterfaces || !rt().itf[itf].live) {
    set_error("%s: no interface %d", who, itf);
    return nullptr;
  }
  return &rt().itf[itf];
}

}  // namespace
}  // namespace bcfd

using namespace bcfd;

extern "C" {

int biocfd_find_dist_node(int64_t n_nodes, const double *xnode, const double *znode, const int64_t *node_id,
                          int64_t markers[3]) {
  BCFD_API_LOCK();
  // the reference keeps the signed coordinate of the last choice as the
  // bound (Search.F90:29-44); only ID-51 nodes are candidates
  double dist = 0.0, dist1 = 999999.0, dist2 = 0.0;
  markers[0] = markers[1] = markers[2] = 0;
  for (int64_t i = 1; i <= n_nodes; ++i) {
    if (node_id[i - 1] != 51) continue;
    if (std::fabs(znode[i - 1]) > dist) { dist = znode[i - 1]; markers[0] = i; }
    if (std::fabs(xnode[i - 1]) < dist1) { dist1 = xnode[i - 1]; markers[1] = i; }
    if (std::fabs(xnode[i - 1]) > dist2) { dist2 = xnode[i - 1]; markers[2] = i; }
  }
  return 0;
}

int biocfd_block_move_check(int itf, const double *xnode1, const double *ynode1, const int64_t markers[3],
                            double *state, int64_t *moved, int64_t *coarse_flcnt_check) {
  BCFD_API_LOCK();
  Interface *it = live_interface(itf, "block_move_check");
  if (!it) return -1;
  Block *b = get_block(it->b);
  Block *c1 = get_block(1);
  if (!b || !c1) return -1;
  *moved = 0;
  for (int q = 0; q < 3; ++q)
    if (markers[q] < 1 || markers[q] > b->ib.n_nodes) {
      set_error("block_move_check: block %d has no ID-51 marker nodes (findDistnode)", it->b);
      return -1;
    }
  const double mg1 = (double)b->cintp * c1->dx;
  const double marginx = 4.5 * mg1, marginy = 5.0 * mg1, marginz = 1.52 * mg1;
  const int64_t factor = it->b_msh / it->a_msh;
  const double ymk = ynode1[markers[0] - 1];
  const double yval_up = dmax(ymk, state[S_NXTY] + 0.02);
  const double yval_dw = dmin(ymk, state[S_NXTY] - 0.02);
  const double xval_lt = xnode1[markers[1] - 1];
  double *e = it->ext;  // xs, xe, ys, ye, zs, ze
  const double pivz = state[S_PIVZ];
  MoveState &mv = b->mv;
  mv.amt[0] = mv.amt[1] = mv.amt[2] = 0;
  if (!(std::fabs(xval_lt - e[0]) <= marginx || std::fabs(pivz - e[4]) <= marginz ||
        std::fabs(pivz - e[5]) <= marginz || std::fabs(yval_dw - e[2]) <= marginy ||
        std::fabs(yval_up - e[3]) <= marginy))
    return 0;
  *moved = 1;
  *coarse_flcnt_check = 1;
  if (std::fabs(yval_dw - e[2]) <= marginy || std::fabs(yval_up - e[3]) <= marginy) {
    mv.amt[1] = whole_coarse((int64_t)std::floor((state[S_NXTY] - state[S_INITY]) / b->dy), factor);
    state[S_INITY] = state[S_NXTY];
  }
  if (std::fabs(xval_lt - e[0]) < marginx) {
    mv.amt[0] = whole_coarse((int64_t)std::floor((state[S_NXTX] - state[S_INITX]) / b->dx) - factor, factor);
    state[S_INITX] = state[S_NXTX];
  }
  if (std::fabs(pivz - e[4]) < (0.500 - 0.410 - mg1) || std::fabs(pivz - e[5]) < (0.600 - 0.500 - mg1)) {
    mv.amt[2] = whole_coarse((int64_t)std::floor((state[S_NXTZ] - state[S_INITZ]) / b->dz), factor);
    state[S_INITZ] = state[S_NXTZ];
  }
  const double h[3] = {b->dx, b->dy, b->dz};
  double *en = it->ext_new;
  for (int a = 0; a < 3; ++a) {
    const double shift = (double)mv.amt[a] * h[a];
    en[2 * a] = dmax(e[2 * a], e[2 * a] + shift);
    en[2 * a + 1] = dmin(e[2 * a + 1], e[2 * a + 1] + shift);
  }
  for (int a = 0; a < 3; ++a) {
    const double shift = (double)mv.amt[a] * h[a];
    e[2 * a] = e[2 * a] + shift;
    e[2 * a + 1] = e[2 * a + 1] + shift;
  }
  copy_range(b->hg[BG_XP], en[0], en[1], &mv.cpy[0], &mv.cpy[1]);
  copy_range(b->hg[BG_YP], en[2], en[3], &mv.cpy[2], &mv.cpy[3]);
  copy_range(b->hg[BG_ZP], en[4], en[5], &mv.cpy[4], &mv.cpy[5]);
  const int fld[4] = {BF_U, BF_V, BF_W, BF_P};
  for (int q = 0; q < 4; ++q) {
    if (!mv.dum[q] && dev_alloc(&mv.dum[q], b->n2)) return -1;
    BCFD_HIP(hipMemcpyAsync(mv.dum[q], b->f[fld[q]], sizeof(double) * b->n2, hipMemcpyDeviceToDevice, rt().stream));
  }
  return 0;
}

int biocfd_change_block_coords(int id, int64_t move_check) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (move_check != 1) return 0;
  const double cz = (double)b->mv.amt[2] * b->dz;
  const double cy = (double)b->mv.amt[1] * b->dy;
  const double cx = (double)b->mv.amt[0] * b->dx;
  std::vector<double> *hg = b->hg;
  const int n[3] = {b->d.nx, b->d.ny, b->d.nz};
  const double c[3] = {cx, cy, cz};
  const int face[3] = {BG_X1, BG_Y1, BG_Z1};
  // faces shifted; u-, v-, w-face coordinates are the faces, the rest
  // mid-points (Search.F90:1218-1262)
  const int same[3] = {BG_XU, BG_YV, BG_ZW};
  const int mids[3][3] = {{BG_XV, BG_XW, BG_XP}, {BG_YU, BG_YW, BG_YP}, {BG_ZU, BG_ZV, BG_ZP}};
  for (int a = 0; a < 3; ++a) {
    std::vector<double> &x1 = hg[face[a]];
    if ((int)x1.size() != n[a] + 3) { set_error("change_block_coords: grids of block %d not uploaded", id); return -1; }
    for (int i = 0; i < n[a] + 3; ++i) x1[i] = x1[i] + c[a];
    hg[same[a]] = x1;
    std::vector<double> mid(n[a] + 2);
    for (int i = 0; i < n[a] + 2; ++i) mid[i] = 0.5 * (x1[i] + x1[i + 1]);
    for (int q = 0; q < 3; ++q) hg[mids[a][q]] = mid;
  }
  const int changed[15] = {BG_X1, BG_Y1, BG_Z1, BG_XU, BG_YU, BG_ZU, BG_XV, BG_YV, BG_ZV,
                           BG_XW, BG_YW, BG_ZW, BG_XP, BG_YP, BG_ZP};
  for (int w : changed)
    BCFD_HIP(hipMemcpyAsync(b->g[w], hg[w].data(), sizeof(double) * hg[w].size(), hipMemcpyHostToDevice,
                            rt().stream));
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  return 0;
}

int biocfd_change_block_coords_interfaces(int itf, int64_t move_check) {
  BCFD_API_LOCK();
  Interface *it = live_interface(itf, "change_block_coords_interfaces");
  if (!it) return -1;
  Block *b = get_block(it->b);
  if (!b) return -1;
  if (move_check != 1) return 0;
  MoveState &mv = b->mv;
  const double *en = it->ext_new;
  hipStream_t s = rt().stream;
  dim3 blk(64, 1, 1), grd(div_up(b->d.NX, 64), b->d.NY, b->d.NZ);
  hipLaunchKernelGGL(k_interior_mask, grd, blk, 0, s, b->d, b->g[BG_XP], b->g[BG_YP], b->g[BG_ZP], en[0] + b->dx,
                     en[1] - b->dx, en[2] + b->dy, en[3] - b->dy, en[4] + b->dz, en[5] - b->dz, b->tag[BT_CELL_N]);
  copy_range(b->hg[BG_XP], en[0], en[1], &mv.cpy_mv[0], &mv.cpy_mv[1]);
  copy_range(b->hg[BG_YP], en[2], en[3], &mv.cpy_mv[2], &mv.cpy_mv[3]);
  copy_range(b->hg[BG_ZP], en[4], en[5], &mv.cpy_mv[4], &mv.cpy_mv[5]);
  Box bx;
  const int N[3] = {b->d.NX, b->d.NY, b->d.NZ};
  bool empty = false;
  for (int a = 0; a < 3; ++a) {
    bx.lo[a] = (int)mv.cpy_mv[2 * a];
    bx.hi[a] = (int)mv.cpy_mv[2 * a + 1];
    bx.src[a] = (int)mv.cpy[2 * a];
    if (bx.hi[a] < bx.lo[a]) empty = true;
  }
  if (empty) {
    for (int a = 0; a < 3; ++a) { bx.lo[a] = 1; bx.hi[a] = 0; bx.src[a] = 1; }
  } else {
    for (int a = 0; a < 3; ++a)
      if (bx.lo[a] < 1 || bx.hi[a] > N[a] || bx.src[a] < 1 || bx.src[a] + (bx.hi[a] - bx.lo[a]) > N[a]) {
        set_error("change_block_coords_interfaces: the kept overlap of interface %d leaves block %d", itf, it->b);
        return -1;
      }
  }
  if (!mv.dum[0]) { set_error("change_block_coords_interfaces: block %d has not been moved", it->b); return -1; }
  hipLaunchKernelGGL(k_shift_fields, grd, blk, 0, s, b->d, bx, mv.dum[0], mv.dum[1], mv.dum[2], mv.dum[3],
                     b->f[BF_U], b->f[BF_V], b->f[BF_W], b->f[BF_P]);
  BCFD_CHECK_LAUNCH();
  return 0;
}

int biocfd_change_block_interface(int itf, int64_t move_check) {
  BCFD_API_LOCK();
  Interface *it = live_interface(itf, "change_block_interface");
  if (!it) return -1;
  Block *b = get_block(it->b);
  if (!b) return -1;
  if (move_check != 1) return 0;
  const int64_t factor = it->b_msh / it->a_msh;
  for (int ax = 0; ax < 3; ++ax) {
    const int64_t inc = b->mv.amt[ax] / factor;
    for (int comp = 0; comp < 4; ++comp) {
      std::vector<int64_t> &h = it->hdet[comp][ax];
      for (int64_t n = 0; n < it->counter[comp][ax]; ++n) h[3 * n] += inc;
      std::vector<int> t32(h.begin(), h.end());
      BCFD_HIP(hipMemcpyAsync(it->det[comp][ax], t32.data(), sizeof(int) * t32.size(), hipMemcpyHostToDevice,
                              rt().stream));
      BCFD_HIP(hipStreamSynchronize(rt().stream));
    }
  }
  return 0;
}

int biocfd_cell_count_solid_coarse_mv(int id, int64_t *coarse_flcnt_check) {
  BCFD_API_LOCK();
  if (*coarse_flcnt_check == 1 && biocfd_cell_count_solid_coarse(id)) return -1;
  *coarse_flcnt_check = 0;
  return 0;
}

int biocfd_tagging_move(int id, int64_t move_check) {
  BCFD_API_LOCK();
  if (move_check != 1) return 0;
  if (biocfd_tagging(id)) return -1;
  for (int q = 1; q <= kMaxInterfaces; ++q) {
    const Interface &it = rt().itf[q];
    if (!it.live || it.b != id) continue;
    if (interface_fine_update_mv(q) || interface_fine_update_one(q, 0)) return -1;
  }
  return 0;
}

int biocfd_grid_download(int id, int which, double *host, int64_t n) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (which < 0 || which >= BG_NGRID || (int64_t)b->hg[which].size() != n) {
    set_error("grid_download: block %d grid %d has %lld entries, not %lld", id, which,
              (long long)(which >= 0 && which < BG_NGRID ? (int64_t)b->hg[which].size() : -1), (long long)n);
    return -1;
  }
  for (int64_t q = 0; q < n; ++q) host[q] = b->hg[which][q];
  return 0;
}

}  // extern "C"
