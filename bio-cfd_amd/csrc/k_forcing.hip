// Sharp-interface IB reconstruction on gfx950 (reference src/Forcing.f90)
// and the body-force sums (src/StressCalc.f90).
//
// Each IB (or TS) cell carries up to six reconstructions ("tasks"): the two
// faces of u, v and w (or one pressure node).  A task sends an image point
// along the surface normal, interpolates value and normal derivative there
// (face-bilinear + linear, Forcing.f90:1091-1161) and fits the quadratic
// profile through the surface value.
//
// The reference runs the tasks serially, writing into the same array it
// interpolates from (velocityForcing1 reads and writes ut; pressureForcing1 and
// pressureForcingField read and write p), so a task may see values written by
// earlier tasks.  That order is reproduced exactly without serialising:
// every task keys itself by its serial position (6*cell + part), each value
// it reads comes from the last earlier task that writes that location (found
// through the cell -> IB-ordinal map) or from the untouched array, and tasks
// run in dependency waves until all are done; results are scattered at the
// end, the last writer of each location winning.
#include "biocfd_rt.hpp"

namespace bcfd {

enum Routine { R_V1 = 0, R_P1, R_VF, R_PF, R_VG, R_PG };

struct FArgs {
  Dims d;
  const double *g[BG_NGRID];
  int64_t n_cells;
  const int *ijk;
  int64_t cap;
  const int *index_ts;     // TS routines: 1-based IB ordinal of each TS cell
  const int *nel[BN_NNEL];
  const double *nd[BN_NNEL];
  const double *el[BE_NELEM];
  const int *surf;
  double kin[BK_N];
  const int *ib_map;
  const double *fld[4];    // arrays interpolated: p, ut, vt, wt
  const double *plain[4];  // arrays copied into the "t" ghosts: p, u, v, w
  int *miss;
};

struct PartSpec {
  int comp;         // 0 p, 1 u, 2 v, 3 w
  int slot;         // BN_*
  int bx, by, bz;   // base point offsets (index shift on the component grid)
  int tx, ty, tz;   // written location offset
  int ghost, ghost_t;
};

__device__ inline PartSpec part_spec(int routine, int part) {
  PartSpec s{};
  if (routine == R_P1 || routine == R_PF || routine == R_PG) {
    s.comp = 0; s.slot = BN_P; s.ghost = BGH_P; s.ghost_t = BGH_PT;
    return s;
  }
  switch (part) {
    case 0: s = {1, BN_U2, 1, 0, 0, 0, 0, 0, BGH_U2, BGH_U2T}; break;
    case 1: s = {1, BN_U1, 0, 0, 0, -1, 0, 0, BGH_U1, BGH_U1T}; break;
    case 2: s = {2, BN_V2, 0, 1, 0, 0, 0, 0, BGH_V2, BGH_V2T}; break;
    case 3: s = {2, BN_V1, 0, 0, 0, 0, -1, 0, BGH_V1, BGH_V1T}; break;
    case 4: s = {3, BN_W2, 0, 0, 1, 0, 0, 0, BGH_W2, BGH_W2T}; break;
    default: s = {3, BN_W1, 0, 0, 0, 0, 0, -1, BGH_W1, BGH_W1T}; break;
  }
  return s;
}

__device__ inline void comp_grids(const FArgs &a, int comp, const double *&gx, const double *&gy,
                                  const double *&gz, int &lx, int &ly, int &lz) {
  const int nx = a.d.nx, ny = a.d.ny, nz = a.d.nz;
  switch (comp) {
    case 0: gx = a.g[BG_XP]; gy = a.g[BG_YP]; gz = a.g[BG_ZP]; lx = nx + 2; ly = ny + 2; lz = nz + 2; break;
    case 1: gx = a.g[BG_XU]; gy = a.g[BG_YU]; gz = a.g[BG_ZU]; lx = nx + 3; ly = ny + 2; lz = nz + 2; break;
    case 2: gx = a.g[BG_XV]; gy = a.g[BG_YV]; gz = a.g[BG_ZV]; lx = nx + 2; ly = ny + 3; lz = nz + 2; break;
    default: gx = a.g[BG_XW]; gy = a.g[BG_YW]; gz = a.g[BG_ZW]; lx = nx + 2; ly = ny + 2; lz = nz + 3; break;
  }
}

// find_index_in_array (Forcing.f90:1165-1183): first i in [s,f] with
// a(i) <= v < a(i+1); the grids are strictly increasing so the hit is unique
__device__ inline int find_index(double v, const double *arr, int len, int s, int f, int *miss) {
  if (s < 1) s = 1;
  if (f > len - 1) f = len - 1;
  if (f - s > 16) {
    int lo = s, hi = f;  // largest i in [s,f] with arr(i) <= v
    if (!(v >= arr[lo - 1])) { atomicAdd(miss, 1); return s; }
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (arr[mid - 1] <= v) lo = mid; else hi = mid - 1;
    }
    if (v >= arr[lo - 1] && v < arr[lo]) return lo;
    atomicAdd(miss, 1);
    return s;
  }
  for (int i = s; i <= f; ++i)
    if (v >= arr[i - 1] && v < arr[i]) return i;
  atomicAdd(miss, 1);
  return s;
}

__device__ inline double bilin(double x, double y, double x1, double x2, double y1, double y2, double f1,
                               double f2, double f3, double f4) {
  const double numer = f1 * (x2 - x) * (y2 - y) + f3 * (x2 - x) * (y - y1) + f2 * (x - x1) * (y2 - y) +
                       f4 * (x - x1) * (y - y1);
  const double denom = (x2 - x1) * (y2 - y1);
  return numer / denom;
}

__device__ inline double lin(double x, double x0, double x1, double y0, double y1) {
  return y0 + (x - x0) * (y1 - y0) / (x1 - x0);
}

__device__ inline double lagrange_d(double x, double x2, double x1, double px, double px2, double px1) {
  const double h2 = fabs(x2 - x), h1 = fabs(x1 - x);
  return (h1 * h1 * px2 - h2 * h2 * px1 + (h2 * h2 - h1 * h1) * px) / (h1 * h2 * (h1 + h2) + 1e-16);
}

// Reads of the interpolated array, serial-order aware for the in-place routines.
struct Reader {
  const FArgs *a;
  int comp;
  bool ordered;
  int routine;
  long long key;           // serial key of the reading task
  const double *task_out;
  const int *task_done;
  int pass;
  bool *blocked;

  __device__ long long writer_before(int64_t loc, int ii, int jj, int kk) const {
    // candidate writers of `loc` for this component and their serial keys
    const Dims &d = a->d;
    long long best = -1;
    if (comp == 0) {
      const int n = a->ib_map[loc];
      if (n >= 0 && (long long)n < key) best = n;
      return best;
    }
    const int p2 = 2 * (comp - 1), p1 = p2 + 1;
    int n = a->ib_map[loc];
    if (n >= 0) { long long kk2 = 6LL * n + p2; if (kk2 < key && kk2 > best) best = kk2; }
    const int si = (comp == 1), sj = (comp == 2), sk = (comp == 3);
    const int ni = ii + si, nj = jj + sj, nk = kk + sk;
    if (ni <= d.NX && nj <= d.NY && nk <= d.NZ) {
      n = a->ib_map[ix3(d, ni, nj, nk)];
      if (n >= 0) { long long kk1 = 6LL * n + p1; if (kk1 < key && kk1 > best) best = kk1; }
    }
    return best;
  }

  __device__ double operator()(int ii, int jj, int kk) const {
    const int64_t loc = ix3(a->d, ii, jj, kk);
    const double *arr = a->fld[comp];
    if (!ordered) return arr[loc];
    const long long w = writer_before(loc, ii, jj, kk);
    if (w < 0) return arr[loc];
    if (task_done[w] >= pass) *blocked = true;
    return task_out[w];
  }
};

// compute_value_and_derivatives (Forcing.f90:1091-1161)
__device__ inline void value_derivs(const Reader &rd, double x, double y, double z, int i, int j, int k,
                                    const double *xg, const double *yg, const double *zg, int off,
                                    double &val, double dv[3]) {
  int ii = i, jj = j, kk = k;
  if (off == 1) ii = i - 1;
  if (off == 2) jj = j - 1;
  if (off == 3) kk = k - 1;
  const double v000 = rd(ii, jj, kk), v100 = rd(ii + 1, jj, kk), v010 = rd(ii, jj + 1, kk);
  const double v110 = rd(ii + 1, jj + 1, kk), v001 = rd(ii, jj, kk + 1), v101 = rd(ii + 1, jj, kk + 1);
  const double v011 = rd(ii, jj + 1, kk + 1), v111 = rd(ii + 1, jj + 1, kk + 1);
  const double X0 = xg[i - 1], X1 = xg[i], Y0 = yg[j - 1], Y1 = yg[j], Z0 = zg[k - 1], Z1 = zg[k];
  const double t11 = bilin(y, z, Y0, Y1, Z0, Z1, v000, v010, v001, v011);
  const double t12 = bilin(y, z, Y0, Y1, Z0, Z1, v100, v110, v101, v111);
  const double t21 = bilin(x, z, X0, X1, Z0, Z1, v000, v100, v001, v101);
  const double t22 = bilin(x, z, X0, X1, Z0, Z1, v010, v110, v011, v111);
  const double t31 = bilin(x, y, X0, X1, Y0, Y1, v000, v100, v010, v110);
  const double t32 = bilin(x, y, X0, X1, Y0, Y1, v001, v101, v011, v111);
  val = lin(x, X0, X1, t11, t12);
  dv[0] = lagrange_d(x, X1, X0, val, t12, t11);
  dv[1] = lagrange_d(y, Y1, Y0, val, t22, t21);
  dv[2] = lagrange_d(z, Z1, Z0, val, t32, t31);
}

// surface value (velocity) or dp/dn (pressure) of element e
__device__ inline double surface_term(const FArgs &a, int routine, int part, int comp, int e) {
  const int sid = a.surf[e];
  const double *k = a.kin;
  const bool flap = (sid == 51 || sid == 52);
  const double td = (sid == 51) ? k[BK_THETADOT1] : k[BK_THETADOT2];
  const double tdd = (sid == 51) ? k[BK_THETADDOT1] : k[BK_THETADDOT2];
  const double xc = a.el[BE_XCENT][e], yc = a.el[BE_YCENT][e], zc = a.el[BE_ZCENT][e];
  if (comp == 0) {
    double ac_z = 0.0, ac_y = 0.0, at_z = 0.0, at_y = 0.0;
    if (flap) {
      ac_z = -(td * td) * (zc - k[BK_PIVZ]);
      ac_y = -(td * td) * (yc - k[BK_PIVY]);
      at_z = tdd * (yc - k[BK_PIVY]);
      at_y = -tdd * (zc - k[BK_PIVZ]);
    }
    const double cb = a.el[BE_COSB][e], cg = a.el[BE_COSG][e];
    if (routine == R_PG) return ((ac_z + at_z) * cg + (ac_y + at_y) * cb) + (k[BK_YDDOT] * cb);
    return -((ac_z + at_z) * cg + (ac_y + at_y) * cb) - k[BK_YDDOT] * cb;
  }
  if (comp == 1) return k[BK_XDOT];
  if (comp == 3) return flap ? td * (yc - k[BK_PIVY]) : 0.0;
  // comp == 2
  const bool alpha = (routine != R_V1);
  if (!flap) {
    const bool plus_zero = (routine == R_V1 && part == 3) || (routine == R_VG && part == 2);
    return plus_zero ? 0.0 + k[BK_YDOT] : k[BK_YDOT];
  }
  if (!alpha) return -td * (zc - k[BK_PIVZ]) + k[BK_YDOT];
  return -td * (zc - k[BK_PIVZ]) + (-k[BK_ALPHADOT]) * (xc - k[BK_PIVX]) + k[BK_YDOT];
}

// one reconstruction; returns false when an ordered read is not available yet
__device__ bool run_task(const FArgs &a, int routine, int64_t cell_ord, int part, const Reader &rd_in,
                         double &out) {
  const Dims &d = a.d;
  const int i = a.ijk[cell_ord], j = a.ijk[cell_ord + a.cap], k = a.ijk[cell_ord + 2 * a.cap];
  const bool ghost = (routine == R_VG || routine == R_PG);
  const int64_t src = ghost ? (int64_t)a.index_ts[cell_ord] - 1 : cell_ord;  // IB ordinal of the data
  const PartSpec s = part_spec(routine, part);
  const double *gx, *gy, *gz;
  int lx, ly, lz;
  comp_grids(a, s.comp, gx, gy, gz, lx, ly, lz);
  const int e = a.nel[s.slot][src] - 1;
  const double sgn = ghost ? -1.0 : 1.0;
  const double sur2 = ghost ? -a.nd[s.slot][src] : a.nd[s.slot][src];
  const double dxi = a.g[BG_DX][i - 1], dyj = a.g[BG_DY][j - 1], dzk = a.g[BG_DZ][k - 1];
  const double pt1 = 1.5 * sqrt(dxi * dxi + dyj * dyj + dzk * dzk) + (fabs(sur2) - sur2) * 0.5;
  const double ca = a.el[BE_COSA][e], cb = a.el[BE_COSB][e], cg = a.el[BE_COSG][e];
  const double bx = gx[i + s.bx - 1], by = gy[j + s.by - 1], bz = gz[k + s.bz - 1];
  double px, py, pz;
  if (sgn > 0) { px = bx + pt1 * ca; py = by + pt1 * cb; pz = bz + pt1 * cg; }
  else { px = bx - pt1 * ca; py = by - pt1 * cb; pz = bz - pt1 * cg; }
  // search windows (Forcing.f90:68-70, 134-341 and the +-7 windows elsewhere)
  int sx0, sx1, sy0, sy1, sz0, sz1;
  if (routine == R_V1 || routine == R_P1) {
    sx0 = sy0 = sz0 = 1;
    sx1 = d.nx + 1; sy1 = d.ny + 1; sz1 = d.nz + 1;
    if (routine == R_V1) {
      if (part == 0) sx1 = d.nx + 2;
      if (part == 2) sy1 = d.ny + 2;
      if (part == 4) sz1 = d.nz + 2;
    }
  } else {
    sx0 = i - 7; sx1 = i + 7; sy0 = j - 7; sy1 = j + 7; sz0 = k - 7; sz1 = k + 7;
  }
  int ix = find_index(px, gx, lx, sx0, sx1, a.miss);
  int iy = find_index(py, gy, ly, sy0, sy1, a.miss);
  int iz = find_index(pz, gz, lz, sz0, sz1, a.miss);
  if (routine == R_V1) {
    if (part == 0 && ix == d.nx + 2) ix = d.nx + 1;
    if (part == 1 && ix == 1) ix = 2;
    if (part == 2 && iy == d.ny + 2) iy = d.ny + 1;
    if (part == 3 && iy == 1) iy = 2;
    if (part == 4 && iz == d.nz + 2) iz = d.nz + 1;
    if (part == 5 && iz == 1) iz = 2;
  }
  // keep the interpolation cube inside the arrays (a miss is reported)
  const int mx = (s.comp == 1) ? 2 : 1, my = (s.comp == 2) ? 2 : 1, mz = (s.comp == 3) ? 2 : 1;
  ix = ix < mx ? mx : (ix > d.NX + mx - 2 ? d.NX + mx - 2 : ix);
  iy = iy < my ? my : (iy > d.NY + my - 2 ? d.NY + my - 2 : iy);
  iz = iz < mz ? mz : (iz > d.NZ + mz - 2 ? d.NZ + mz - 2 : iz);
  bool blocked = false;
  Reader rd = rd_in;
  rd.comp = s.comp;
  rd.blocked = &blocked;
  double val, dv[3];
  value_derivs(rd, px, py, pz, ix, iy, iz, gx, gy, gz, s.comp, val, dv);
  if (blocked) return false;
  double dn = dv[0] * ca + dv[1] * cb + dv[2] * cg;
  if (ghost) dn = -1.0 * dn;
  const double n1 = pt1 + sur2;
  const double surfv = surface_term(a, routine, part, s.comp, e);
  if (s.comp == 0) {
    const double bval = surfv;
    const double aval = (dn - surfv) / (2.0 * n1);
    const double cval = val - (dn + surfv) * n1 * 0.5;
    out = aval * (sur2 * sur2) + bval * sur2 + cval;
  } else {
    const double cval = surfv;
    const double bval = 2.0 / n1 * (val - surfv) - dn;
    const double aval = dn / n1 - (val - surfv) / (n1 * n1);
    out = aval * (sur2 * sur2) + bval * sur2 + cval;
  }
  return true;
}

// ordered routines: one dependency wave
__global__ void k_force_pass(FArgs a, int routine, int nparts, double *task_out, int *task_done, int pass,
                             int *remaining) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_cells * nparts) return;
  if (task_done[t] < pass) return;  // finished in an earlier wave
  Reader rd{&a, 0, true, routine, (long long)t, task_out, task_done, pass, nullptr};
  double out;
  if (run_task(a, routine, t / nparts, (int)(t % nparts), rd, out)) {
    task_out[t] = out;
    task_done[t] = pass;
  } else {
    atomicAdd(remaining, 1);
  }
}

__global__ void k_force_init(int *task_done, int64_t n) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) task_done[t] = 0x7fffffff;
}

// scatter: the last writer of each location wins
__global__ void k_force_scatter(FArgs a, int routine, int nparts, const double *task_out, double *f0, double *f1,
                                double *f2, double *f3) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_cells * nparts) return;
  const int64_t n = t / nparts;
  const int part = (int)(t % nparts);
  const PartSpec s = part_spec(routine, part);
  const int i = a.ijk[n] + s.tx, j = a.ijk[n + a.cap] + s.ty, k = a.ijk[n + 2 * a.cap] + s.tz;
  const Dims &d = a.d;
  if (routine == R_V1) {
    // the other writer of this location is the neighbouring IB cell's opposite face
    const bool first = (part % 2) == 0;
    const int si = (s.comp == 1), sj = (s.comp == 2), sk = (s.comp == 3);
    long long other = -1;
    if (first) {
      const int ni = i + si, nj = j + sj, nk = k + sk;
      if (ni <= d.NX && nj <= d.NY && nk <= d.NZ) {
        const int m = a.ib_map[ix3(d, ni, nj, nk)];
        if (m >= 0) other = 6LL * m + part + 1;
      }
    } else {
      const int m = a.ib_map[ix3(d, i, j, k)];
      if (m >= 0) other = 6LL * m + part - 1;
    }
    if (other > t) return;
  }
  double *dst = (s.comp == 0) ? f0 : (s.comp == 1) ? f1 : (s.comp == 2) ? f2 : f3;
  dst[ix3(d, i, j, k)] = task_out[t];
}

// routines without in-place reads: direct
__global__ void k_force_direct(FArgs a, int routine, int nparts, double *f0, double *f1, double *f2, double *f3,
                               double *const *ghost) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_cells * nparts) return;
  const int64_t n = t / nparts;
  const int part = (int)(t % nparts);
  Reader rd{&a, 0, false, routine, 0, nullptr, nullptr, 0, nullptr};
  double out = 0.0;
  run_task(a, routine, n, part, rd, out);
  const PartSpec s = part_spec(routine, part);
  const int i = a.ijk[n] + s.tx, j = a.ijk[n + a.cap] + s.ty, k = a.ijk[n + 2 * a.cap] + s.tz;
  const int64_t loc = ix3(a.d, i, j, k);
  if (routine == R_VF) {
    double *dst = (s.comp == 1) ? f1 : (s.comp == 2) ? f2 : f3;
    dst[loc] = out;
  } else {
    ghost[s.ghost][n] = out;
    ghost[s.ghost_t][n] = a.plain[s.comp][loc];
  }
}

static int force(int id, int routine, const double *kin) {
  Block *b = get_block(id);
  if (!b) return -1;
  IbData &ib = b->ib;
  if (!ib.present) { set_error("block %d has no immersed body", id); return -1; }
  const bool ts = (routine == R_VG || routine == R_PG);
  const int64_t n_cells = ts ? b->counts[BC_TS] : b->counts[BC_IB];
  if (n_cells == 0) return 0;
  if (!ts && ib.n_mapped != n_cells) {
    set_error("forcing: IB list is stale (cell_count_solid after retagging is required)");
    return -1;
  }
  const int nparts = (routine == R_P1 || routine == R_PF || routine == R_PG) ? 1 : 6;
  hipStream_t s = rt().stream;
  FArgs a{};
  a.d = b->d;
  for (int q = 0; q < BG_NGRID; ++q) {
    if (!b->g[q]) { set_error("grid %d not uploaded", q); return -1; }
    a.g[q] = b->g[q];
  }
  a.n_cells = n_cells;
  a.ijk = ts ? ib.ts_ijk : ib.ib_ijk;
  a.cap = ts ? ib.cap_ts : ib.cap_ib;
  a.index_ts = ib.index_ts;
  for (int q = 0; q < BN_NNEL; ++q) { a.nel[q] = ib.nel[q]; a.nd[q] = ib.normdis[q]; }
  for (int q = 0; q < BE_NELEM; ++q) a.el[q] = ib.elem[q];
  a.surf = ib.surf;
  for (int q = 0; q < BK_N; ++q) a.kin[q] = kin[q];
  a.ib_map = ib.ib_map;
  a.fld[0] = b->f[BF_P]; a.fld[1] = b->f[BF_UT]; a.fld[2] = b->f[BF_VT]; a.fld[3] = b->f[BF_WT];
  a.plain[0] = b->f[BF_P]; a.plain[1] = b->f[BF_U]; a.plain[2] = b->f[BF_V]; a.plain[3] = b->f[BF_W];
  const int64_t ntask = n_cells * nparts;
  if (grow(&ib.task_out, &ib.cap_task, ntask + 2)) return -1;
  if (grow(&ib.task_done, &ib.cap_done, ntask + 2)) return -1;
  int *counters = (int *)(ib.task_out + ib.cap_task - 1);  // 2 ints in the last slot
  a.miss = counters;
  BCFD_HIP(hipMemsetAsync(counters, 0, 2 * sizeof(int), s));
  const unsigned g = div_up(ntask, 256);
  if (routine == R_VF || ts) {
    double **gh = nullptr;
    if (ts) {
      BCFD_HIP(hipMalloc((void **)&gh, sizeof(double *) * BGH_N));
      BCFD_HIP(hipMemcpyAsync(gh, ib.ghost, sizeof(double *) * BGH_N, hipMemcpyHostToDevice, s));
    }
    hipLaunchKernelGGL(k_force_direct, dim3(g), dim3(256), 0, s, a, routine, nparts, b->f[BF_P], b->f[BF_U],
                       b->f[BF_V], b->f[BF_W], (double *const *)gh);
    BCFD_CHECK_LAUNCH();
    BCFD_HIP(hipStreamSynchronize(s));
    if (gh) (void)hipFree(gh);
  } else {
    hipLaunchKernelGGL(k_force_init, dim3(g), dim3(256), 0, s, ib.task_done, ntask);
    int *h = (int *)rt().h_pinned_i;
    int pass = 1;
    for (;; ++pass) {
      BCFD_HIP(hipMemsetAsync(counters + 1, 0, sizeof(int), s));
      hipLaunchKernelGGL(k_force_pass, dim3(g), dim3(256), 0, s, a, routine, nparts, ib.task_out, ib.task_done,
                         pass, counters + 1);
      BCFD_CHECK_LAUNCH();
      BCFD_HIP(hipMemcpyAsync(h, counters, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
      BCFD_HIP(hipStreamSynchronize(s));
      if (h[1] == 0) break;
      if (pass > ntask + 1) { set_error("forcing: dependency waves did not finish"); return -1; }
    }
    double *f0 = b->f[BF_P];
    double *f1 = b->f[BF_UT], *f2 = b->f[BF_VT], *f3 = b->f[BF_WT];
    hipLaunchKernelGGL(k_force_scatter, dim3(g), dim3(256), 0, s, a, routine, nparts, ib.task_out, f0, f1, f2, f3);
    BCFD_CHECK_LAUNCH();
  }
  int *h = (int *)rt().h_pinned_i;
  BCFD_HIP(hipMemcpyAsync(h, counters, sizeof(int), hipMemcpyDeviceToHost, s));
  BCFD_HIP(hipStreamSynchronize(s));
  if (h[0] != 0) {
    set_error("forcing: %d image-point lookups fell outside their search window", h[0]);
    return -2;
  }
  return 0;
}

// ---------------- stress_calculation (StressCalc.f90:14-208) ----------------
__global__ void k_stress_terms(FArgs a, double mu_f, double rho_f, double *terms) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ne = a.n_cells;
  if (e >= ne) return;
  const Dims &d = a.d;
  const double nrm[3] = {a.el[BE_COSA][e], a.el[BE_COSB][e], a.el[BE_COSG][e]};
  const double xc = a.el[BE_XCENT][e], yc = a.el[BE_YCENT][e], zc = a.el[BE_ZCENT][e];
  const int ic = find_index(xc, a.g[BG_X1], d.nx + 3, 2, d.nx + 1, a.miss);
  const int jc = find_index(yc, a.g[BG_Y1], d.ny + 3, 2, d.ny + 1, a.miss);
  const int kc = find_index(zc, a.g[BG_Z1], d.nz + 3, 2, d.nz + 1, a.miss);
  const double dX = a.g[BG_X1][ic] - a.g[BG_X1][ic - 1];
  const double dY = a.g[BG_Y1][jc] - a.g[BG_Y1][jc - 1];
  const double dZ = a.g[BG_Z1][kc] - a.g[BG_Z1][kc - 1];
  const double diag = sqrt(dX * dX + dY * dY + dZ * dZ);
  const double px = xc + diag * nrm[0], py = yc + diag * nrm[1], pz = zc + diag * nrm[2];
  const int sid = a.surf[e];
  const double *k = a.kin;
  double td = 0.0, tdd = 0.0;
  if (sid == 51) { td = k[BK_THETADOT1]; tdd = k[BK_THETADDOT1]; }
  else if (sid == 52) { td = k[BK_THETADOT2]; tdd = k[BK_THETADDOT2]; }
  const double surf[3] = {k[BK_XDOT], -td * (zc - k[BK_PIVZ]) + k[BK_YDOT], td * (yc - k[BK_PIVY])};
  const double ac_z = -(td * td) * (zc - k[BK_PIVZ]), ac_y = -(td * td) * (yc - k[BK_PIVY]);
  const double at_z = tdd * (yc - k[BK_PIVY]), at_y = -tdd * (zc - k[BK_PIVZ]);
  const double dpdn = -((ac_z + at_z) * nrm[2] + (ac_y + at_y) * nrm[1]) - k[BK_YDDOT] * nrm[1];
  double ddn[3];
  for (int c = 0; c < 3; ++c) {
    const double *gx, *gy, *gz;
    int lx, ly, lz;
    comp_grids(a, c + 1, gx, gy, gz, lx, ly, lz);
    const int ix = find_index(px, gx, lx, 2, d.nx + 1, a.miss);
    const int iy = find_index(py, gy, ly, 2, d.ny + 1, a.miss);
    const int iz = find_index(pz, gz, lz, 2, d.nz + 1, a.miss);
    Reader rd{&a, c + 1, false, 0, 0, nullptr, nullptr, 0, nullptr};
    double val, dv[3];
    value_derivs(rd, px, py, pz, ix, iy, iz, gx, gy, gz, c + 1, val, dv);
    const double dn = dv[0] * nrm[0] + dv[1] * nrm[1] + dv[2] * nrm[2];
    ddn[c] = (2.0 / diag) * (val - surf[c]) - dn;
  }
  const double alen = a.el[BE_ELEN][e];
  const double area = alen / 2.0;
  const double area_yz = 0.5 * fabs(alen * nrm[0]);
  const double area_xz = 0.5 * fabs(alen * nrm[1]);
  const double dd = ddn[0] * nrm[0] + ddn[1] * nrm[1] + ddn[2] * nrm[2];
  const double sh0 = (ddn[0] - dd * nrm[0]) * mu_f * area;
  const double sh1 = (ddn[1] - dd * nrm[1]) * mu_f * area;
  const double *gx, *gy, *gz;
  int lx, ly, lz;
  comp_grids(a, 0, gx, gy, gz, lx, ly, lz);
  const int ix = find_index(px, gx, lx, 2, d.nx + 1, a.miss);
  const int iy = find_index(py, gy, ly, 2, d.ny + 1, a.miss);
  const int iz = find_index(pz, gz, lz, 2, d.nz + 1, a.miss);
  Reader rd{&a, 0, false, 0, 0, nullptr, nullptr, 0, nullptr};
  double val, dv[3];
  value_derivs(rd, px, py, pz, ix, iy, iz, gx, gy, gz, 0, val, dv);
  const double dpdn_e = dv[0] * nrm[0] + dv[1] * nrm[1] + dv[2] * nrm[2];
  const double cval = val - (dpdn_e + dpdn) * diag * 0.5;
  const double f0 = -cval * area * nrm[0] * rho_f;
  const double f1 = -cval * area * nrm[1] * rho_f;
  double *t = terms + 6 * e;
  t[0] = sh0; t[1] = f0; t[2] = sh1; t[3] = f1; t[4] = area_xz; t[5] = area_yz;
}

// The reference sums element by element (StressCalc.f90:75, serial in the
// CPU build).  One workgroup keeps that order: the terms pass through two LDS
// buffers of kStressChunk elements; while waves 1-3 load chunk c+1 into one,
// lanes 0..5 of wave 0 add their term of chunk c's elements, in element
// order, from the other (64 values at a time in registers, so the LDS reads
// are in flight together and only the additions form a chain).
constexpr int kStressChunk = 512;
constexpr int kStressThreads = 256;
__global__ void __launch_bounds__(kStressThreads) k_stress_sum(const double *__restrict__ terms, int64_t ne,
                                                               double *__restrict__ out) {
  __shared__ double buf[2][6 * kStressChunk];
  __shared__ double sums[6];
  const int t = threadIdx.x;
  const int64_t nchunk = (ne + kStressChunk - 1) / kStressChunk;
  auto load = [&](int64_t c, double *dst, int first, int stride) {
    const int64_t e0 = c * kStressChunk;
    const int n = ne - e0 < kStressChunk ? (int)(ne - e0) : kStressChunk;
    for (int q = first; q < 6 * n; q += stride) dst[q] = terms[6 * e0 + q];
  };
  if (nchunk > 0) load(0, buf[0], t, kStressThreads);
  __syncthreads();
  double acc = 0.0;
  for (int64_t c = 0; c < nchunk; ++c) {
    if (t >= 64) {
      if (c + 1 < nchunk) load(c + 1, buf[(c + 1) & 1], t - 64, kStressThreads - 64);
    } else if (t < 6) {
      const double *b = buf[c & 1];
      const int64_t e0 = c * kStressChunk;
      const int n = ne - e0 < kStressChunk ? (int)(ne - e0) : kStressChunk;
      for (int g = 0; g < n; g += 64) {
        double v[64];
#pragma unroll
        for (int q = 0; q < 64; ++q) v[q] = g + q < n ? b[6 * (g + q) + t] : 0.0;
#pragma unroll
        for (int q = 0; q < 64; ++q)
          if (g + q < n) acc = acc + v[q];
      }
    }
    __syncthreads();
  }
  if (t < 6) sums[t] = acc;
  __syncthreads();
  if (t == 0) {
    const double ax = 0.5 * sums[4], ay = 0.5 * sums[5];
    out[0] = 2.0 * (sums[0] / ax);
    out[1] = 2.0 * (sums[1] / ay);
    out[2] = 2.0 * (sums[2] / ax);
    out[3] = 2.0 * (sums[3] / ay);
  }
}

__global__ void k_stress_clear(int *miss) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *miss = 0;
}

// stress_calculation of block b with its terms (6 per element), the four
// coefficients and the miss counter in `work` (6 E + 8 doubles); leaves
// the result in h_pinned
static int stress_enqueue(Block &b, double mu_f, double rho_f, const double kin[BK_N], double *work) {
  IbData &ib = b.ib;
  hipStream_t s = rt().stream;
  FArgs a{};
  a.d = b.d;
  for (int q = 0; q < BG_NGRID; ++q) a.g[q] = b.g[q];
  a.n_cells = ib.n_elems;
  for (int q = 0; q < BE_NELEM; ++q) a.el[q] = ib.elem[q];
  a.surf = ib.surf;
  for (int q = 0; q < BK_N; ++q) a.kin[q] = kin[q];
  a.fld[0] = b.f[BF_P]; a.fld[1] = b.f[BF_U]; a.fld[2] = b.f[BF_V]; a.fld[3] = b.f[BF_W];
  double *res = work + 6 * ib.n_elems;
  a.miss = (int *)(res + 4);
  hipLaunchKernelGGL(k_stress_clear, dim3(1), dim3(64), 0, s, a.miss);
  hipLaunchKernelGGL(k_stress_terms, dim3(div_up(ib.n_elems, 256)), dim3(256), 0, s, a, mu_f, rho_f, work);
  hipLaunchKernelGGL(k_stress_sum, dim3(1), dim3(kStressThreads), 0, s, work, ib.n_elems, res);
  BCFD_CHECK_LAUNCH();
  BCFD_HIP(hipMemcpyAsync(rt().h_pinned, res, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
  BCFD_HIP(hipStreamSynchronize(s));
  return 0;
}

// Scratch of biocfd_stress_calculation_noalloc: the per-element terms (6 per
// element, up to kNoallocElems elements), then the four coefficients and the
// miss counter.  A file-scope device array: each device the library runs on
// holds its own copy (3 MB) until the process ends.  Host code reaches it
// only through the address hipGetSymbolAddress returns for the device.
constexpr int64_t kNoallocElems = 65536;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[6 * kNoallocElems + 8];
static double *g_scratch_ptr[kMaxDevices] = {};

// the scratch address on `device` (current), once per device; called by
// biocfd_init
int stress_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("stress scratch: device %d outside 0..%d", device, kMaxDevices - 1);
    return -1;
  }
  if (g_scratch_ptr[device]) return 0;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error("hipGetSymbolAddress(stress scratch) on device %d: %s", device, hipGetErrorString(e));
    return -1;
  }
  g_scratch_ptr[device] = (double *)p;
  return 0;
}

static int stress_result(double out[4]) {
  const double *h = rt().h_pinned;
  int miss;
  memcpy(&miss, &h[4], sizeof(int));
  for (int q = 0; q < 4; ++q) out[q] = h[q];
  if (miss != 0) { set_error("stress_calculation: %d probe lookups missed", miss); return -2; }
  return 0;
}

}  // namespace bcfd

using namespace bcfd;

extern "C" {

int biocfd_velocity_forcing1(int id, const double kin[BK_N]) {
  BCFD_API_LOCK(); return force(id, R_V1, kin); }
int biocfd_pressure_forcing1(int id, const double kin[BK_N]) {
  BCFD_API_LOCK(); return force(id, R_P1, kin); }
int biocfd_velocity_forcing_field(int id, const double kin[BK_N]) {
  BCFD_API_LOCK(); return force(id, R_VF, kin); }
int biocfd_pressure_forcing_field(int id, const double kin[BK_N]) {
  BCFD_API_LOCK(); return force(id, R_PF, kin); }
int biocfd_velocity_forcing_ghost(int id, const double kin[BK_N]) {
  BCFD_API_LOCK(); return force(id, R_VG, kin); }
int biocfd_pressure_forcing_ghost(int id, const double kin[BK_N]) {
  BCFD_API_LOCK(); return force(id, R_PG, kin); }

int biocfd_stress_calculation(int id, double mu_f, double rho_f, const double kin[BK_N], double out[4]) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  IbData &ib = b->ib;
  if (!ib.present) { set_error("block %d has no immersed body", id); return -1; }
  // per-element terms and the result, kept on the block between calls
  if (grow(&ib.stress_work, &ib.cap_stress, 6 * ib.n_elems + 8)) return -1;
  if (stress_enqueue(*b, mu_f, rho_f, kin, ib.stress_work)) return -1;
  return stress_result(out);
}

// Calls on one device share its scratch: they are serialised by the library
// lock and run in order on the library's stream; calls that overlapped on one
// device (two processes, or another stream) would race on it.
int biocfd_stress_calculation_noalloc(int id, double mu_f, double rho_f, const double kin[BK_N], double out[4]) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  IbData &ib = b->ib;
  if (!ib.present) { set_error("block %d has no immersed body", id); return -1; }
  if (ib.n_elems > kNoallocElems) {
    set_error("stress_calculation_noalloc: %lld elements exceed the scratch maximum of %lld",
              (long long)ib.n_elems, (long long)kNoallocElems);
    return -1;
  }
  const int dev = rt().device;
  double *work = dev >= 0 && dev < kMaxDevices ? g_scratch_ptr[dev] : nullptr;
  if (!work) { set_error("stress_calculation_noalloc: no scratch on device %d (biocfd_init)", dev); return -1; }
  if (stress_enqueue(*b, mu_f, rho_f, kin, work)) return -1;
  return stress_result(out);
}

}  // extern "C"
