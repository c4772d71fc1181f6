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
#include "dual.hpp"

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
  // tangent-linear mode: the surface values are the tangents of the surface
  // terms along kin_dot (the fields passed are tangents too)
  int tangent;
  double kin_dot[BK_N];
  const double *addend;   // tangent mode: per-task geometric tangent added to the output (or null)
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

// The interpolation helpers are generic in their scalar type (double, or
// Dual for the derivatives with respect to the image point and the values).
template <class T>
__device__ inline T bilin(T x, T y, double x1, double x2, double y1, double y2, T f1, T f2, T f3, T f4) {
  const T numer = f1 * (x2 - x) * (y2 - y) + f3 * (x2 - x) * (y - y1) + f2 * (x - x1) * (y2 - y) +
                  f4 * (x - x1) * (y - y1);
  const double denom = (x2 - x1) * (y2 - y1);
  return numer / denom;
}

template <class T>
__device__ inline T lin(T x, double x0, double x1, T y0, T y1) {
  return y0 + (x - x0) * (y1 - y0) / (x1 - x0);
}

template <class T>
__device__ inline T lagrange_d(T x, double x2, double x1, T px, T px2, T px1) {
  const T h2 = fabs(x2 - x), h1 = fabs(x1 - x);
  return (h1 * h1 * px2 - h2 * h2 * px1 + (h2 * h2 - h1 * h1) * px) / (h1 * h2 * (h1 + h2) + 1e-16);
}

template <class T> __device__ inline T as(double x);
template <> __device__ inline double as<double>(double x) { return x; }
template <> __device__ inline Dual as<Dual>(double x) { return {x, 0.0}; }
template <class T> __device__ inline T as(Dual x) { return x; }

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

// compute_value_and_derivatives (Forcing.f90:1091-1161); the reader rd
// returns the cube's values (double, or Dual where their derivatives are
// wanted)
template <class R, class T = double>
__device__ inline void value_derivs(const R &rd, T x, T y, T z, int i, int j, int k, const double *xg,
                                    const double *yg, const double *zg, int off, T &val, T dv[3]) {
  int ii = i, jj = j, kk = k;
  if (off == 1) ii = i - 1;
  if (off == 2) jj = j - 1;
  if (off == 3) kk = k - 1;
  const T v000 = as<T>(rd(ii, jj, kk)), v100 = as<T>(rd(ii + 1, jj, kk)), v010 = as<T>(rd(ii, jj + 1, kk));
  const T v110 = as<T>(rd(ii + 1, jj + 1, kk)), v001 = as<T>(rd(ii, jj, kk + 1));
  const T v101 = as<T>(rd(ii + 1, jj, kk + 1)), v011 = as<T>(rd(ii, jj + 1, kk + 1));
  const T v111 = as<T>(rd(ii + 1, jj + 1, kk + 1));
  const double X0 = xg[i - 1], X1 = xg[i], Y0 = yg[j - 1], Y1 = yg[j], Z0 = zg[k - 1], Z1 = zg[k];
  const T t11 = bilin(y, z, Y0, Y1, Z0, Z1, v000, v010, v001, v011);
  const T t12 = bilin(y, z, Y0, Y1, Z0, Z1, v100, v110, v101, v111);
  const T t21 = bilin(x, z, X0, X1, Z0, Z1, v000, v100, v001, v101);
  const T t22 = bilin(x, z, X0, X1, Z0, Z1, v010, v110, v011, v111);
  const T t31 = bilin(x, y, X0, X1, Y0, Y1, v000, v100, v010, v110);
  const T t32 = bilin(x, y, X0, X1, Y0, Y1, v001, v101, v011, v111);
  val = lin(x, X0, X1, t11, t12);
  dv[0] = lagrange_d(x, X1, X0, val, t12, t11);
  dv[1] = lagrange_d(y, Y1, Y0, val, t22, t21);
  dv[2] = lagrange_d(z, Z1, Z0, val, t32, t31);
}

// surface value (velocity) or dp/dn (pressure) of an element of surface
// sid with centroid c and normal nrm, from the kinematic vector k (generic:
// the adjoint passes differentiate it in c, nrm and k)
template <class T>
__device__ inline T surface_term_T(int routine, int part, int comp, int sid, const T c[3], const T nrm[3],
                                   const T k[BK_N]) {
  const bool flap = (sid == 51 || sid == 52);
  const T td = (sid == 51) ? k[BK_THETADOT1] : k[BK_THETADOT2];
  const T tdd = (sid == 51) ? k[BK_THETADDOT1] : k[BK_THETADDOT2];
  const T xc = c[0], yc = c[1], zc = c[2];
  if (comp == 0) {
    T ac_z = as<T>(0.0), ac_y = as<T>(0.0), at_z = as<T>(0.0), at_y = as<T>(0.0);
    if (flap) {
      ac_z = -(td * td) * (zc - k[BK_PIVZ]);
      ac_y = -(td * td) * (yc - k[BK_PIVY]);
      at_z = tdd * (yc - k[BK_PIVY]);
      at_y = -tdd * (zc - k[BK_PIVZ]);
    }
    const T cb = nrm[1], cg = nrm[2];
    if (routine == R_PG) return ((ac_z + at_z) * cg + (ac_y + at_y) * cb) + (k[BK_YDDOT] * cb);
    return -((ac_z + at_z) * cg + (ac_y + at_y) * cb) - k[BK_YDDOT] * cb;
  }
  if (comp == 1) return k[BK_XDOT];
  if (comp == 3) return flap ? td * (yc - k[BK_PIVY]) : as<T>(0.0);
  // comp == 2
  const bool alpha = (routine != R_V1);
  if (!flap) {
    const bool plus_zero = (routine == R_V1 && part == 3) || (routine == R_VG && part == 2);
    return plus_zero ? 0.0 + k[BK_YDOT] : k[BK_YDOT];
  }
  if (!alpha) return -td * (zc - k[BK_PIVZ]) + k[BK_YDOT];
  return -td * (zc - k[BK_PIVZ]) + (-k[BK_ALPHADOT]) * (xc - k[BK_PIVX]) + k[BK_YDOT];
}

__device__ inline double surface_term(const FArgs &a, int routine, int part, int comp, int e) {
  const double c[3] = {a.el[BE_XCENT][e], a.el[BE_YCENT][e], a.el[BE_ZCENT][e]};
  const double nrm[3] = {a.el[BE_COSA][e], a.el[BE_COSB][e], a.el[BE_COSG][e]};
  return surface_term_T<double>(routine, part, comp, a.surf[e], c, nrm, a.kin);
}

// gradient of surface_term with respect to the kinematic vector (the
// tangent-linear and adjoint passes; element positions held fixed)
__device__ inline void surface_term_grad(const FArgs &a, int routine, int comp, int e, double g[BK_N]) {
  for (int q = 0; q < BK_N; ++q) g[q] = 0.0;
  const int sid = a.surf[e];
  const double *k = a.kin;
  const bool flap = (sid == 51 || sid == 52);
  const int std_ = (sid == 51) ? BK_THETADOT1 : BK_THETADOT2;
  const int stdd = (sid == 51) ? BK_THETADDOT1 : BK_THETADDOT2;
  const double td = k[std_], tdd = k[stdd];
  const double xc = a.el[BE_XCENT][e], yc = a.el[BE_YCENT][e], zc = a.el[BE_ZCENT][e];
  const double ry = yc - k[BK_PIVY], rz = zc - k[BK_PIVZ];
  if (comp == 0) {
    const double sg = (routine == R_PG) ? 1.0 : -1.0;
    const double cb = a.el[BE_COSB][e], cg = a.el[BE_COSG][e];
    // S = sg ((ac_z + at_z) cg + (ac_y + at_y) cb + yddot cb)
    g[BK_YDDOT] = sg * cb;
    if (flap) {
      g[std_] = sg * (-2.0 * td * rz * cg - 2.0 * td * ry * cb);
      g[stdd] = sg * (ry * cg - rz * cb);
      g[BK_PIVZ] = sg * (td * td * cg + tdd * cb);
      g[BK_PIVY] = sg * (-tdd * cg + td * td * cb);
    }
    return;
  }
  if (comp == 1) { g[BK_XDOT] = 1.0; return; }
  if (comp == 3) {
    if (flap) { g[std_] = ry; g[BK_PIVY] = -td; }
    return;
  }
  g[BK_YDOT] = 1.0;
  if (!flap) return;
  g[std_] = -rz;
  g[BK_PIVZ] = td;
  if (routine != R_V1) {
    g[BK_ALPHADOT] = -(xc - k[BK_PIVX]);
    g[BK_PIVX] = k[BK_ALPHADOT];
  }
}

// The geometry of one reconstruction: the written component and slot, the
// image point and its interpolation cube, the normal and the distances.
struct TaskGeom {
  PartSpec s;
  const double *gx, *gy, *gz;
  int ix, iy, iz, e;
  double px, py, pz, ca, cb, cg, sur2, n1;
  bool ghost;
};

__device__ inline void task_geom(const FArgs &a, int routine, int64_t cell_ord, int part, TaskGeom &t) {
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
  t.ix = ix < mx ? mx : (ix > d.NX + mx - 2 ? d.NX + mx - 2 : ix);
  t.iy = iy < my ? my : (iy > d.NY + my - 2 ? d.NY + my - 2 : iy);
  t.iz = iz < mz ? mz : (iz > d.NZ + mz - 2 ? d.NZ + mz - 2 : iz);
  t.s = s;
  t.gx = gx; t.gy = gy; t.gz = gz;
  t.e = e;
  t.px = px; t.py = py; t.pz = pz;
  t.ca = ca; t.cb = cb; t.cg = cg;
  t.sur2 = sur2;
  t.n1 = pt1 + sur2;
  t.ghost = ghost;
}

// the quadratic profile's value at the node from the image-point value and
// derivatives and the surface term
template <class T>
__device__ inline T profile_T(bool ghost, int comp, T ca, T cb, T cg, T sur2, T n1, T val, const T dv[3], T surfv) {
  T dn = dv[0] * ca + dv[1] * cb + dv[2] * cg;
  if (ghost) dn = -1.0 * dn;
  if (comp == 0) {
    const T bval = surfv;
    const T aval = (dn - surfv) / (2.0 * n1);
    const T cval = val - (dn + surfv) * n1 * 0.5;
    return aval * (sur2 * sur2) + bval * sur2 + cval;
  }
  const T cval = surfv;
  const T bval = 2.0 / n1 * (val - surfv) - dn;
  const T aval = dn / n1 - (val - surfv) / (n1 * n1);
  return aval * (sur2 * sur2) + bval * sur2 + cval;
}

__device__ inline double task_value(const TaskGeom &t, double val, const double dv[3], double surfv) {
  return profile_T<double>(t.ghost, t.s.comp, t.ca, t.cb, t.cg, t.sur2, t.n1, val, dv, surfv);
}

// One reconstruction as a function of its element's centroid c and normal
// nrm, the kinematic vector kin and the values the reader returns, with the
// forward's image-point cube (g): the normal distance from the slot's node
// (k_norm_distance's expression), the image point, the value and derivatives
// there, the profile.  T = double reproduces run_task; Dual carries one
// derivative.
template <class T, class R>
__device__ inline T task_eval(const FArgs &a, int routine, int part, int64_t cell_ord, const TaskGeom &g,
                              const T c[3], const T nrm[3], const T kin[BK_N], const R &rd) {
  const int i = a.ijk[cell_ord], j = a.ijk[cell_ord + a.cap], k = a.ijk[cell_ord + 2 * a.cap];
  const int slot = g.s.slot;
  const double ptx = slot == BN_U1 ? a.g[BG_X1][i - 1] : (slot == BN_U2 ? a.g[BG_X1][i] : a.g[BG_XP][i - 1]);
  const double pty = slot == BN_V1 ? a.g[BG_Y1][j - 1] : (slot == BN_V2 ? a.g[BG_Y1][j] : a.g[BG_YP][j - 1]);
  const double ptz = slot == BN_W1 ? a.g[BG_Z1][k - 1] : (slot == BN_W2 ? a.g[BG_Z1][k] : a.g[BG_ZP][k - 1]);
  const T nd = (ptx - c[0]) * nrm[0] + (pty - c[1]) * nrm[1] + (ptz - c[2]) * nrm[2];
  const T sur2 = g.ghost ? -nd : nd;
  const double dxi = a.g[BG_DX][i - 1], dyj = a.g[BG_DY][j - 1], dzk = a.g[BG_DZ][k - 1];
  const T pt1 = 1.5 * sqrt(dxi * dxi + dyj * dyj + dzk * dzk) + (fabs(sur2) - sur2) * 0.5;
  const PartSpec &s = g.s;
  const double bx = g.gx[i + s.bx - 1], by = g.gy[j + s.by - 1], bz = g.gz[k + s.bz - 1];
  T px, py, pz;
  if (!g.ghost) { px = bx + pt1 * nrm[0]; py = by + pt1 * nrm[1]; pz = bz + pt1 * nrm[2]; }
  else { px = bx - pt1 * nrm[0]; py = by - pt1 * nrm[1]; pz = bz - pt1 * nrm[2]; }
  T val, dv[3];
  value_derivs<R, T>(rd, px, py, pz, g.ix, g.iy, g.iz, g.gx, g.gy, g.gz, s.comp, val, dv);
  const T n1 = pt1 + sur2;
  const T surfv = surface_term_T<T>(routine, part, s.comp, a.surf[g.e], c, nrm, kin);
  return profile_T<T>(g.ghost, s.comp, nrm[0], nrm[1], nrm[2], sur2, n1, val, dv, surfv);
}

// one reconstruction; returns false when an ordered read is not available yet
__device__ bool run_task(const FArgs &a, int routine, int64_t cell_ord, int part, const Reader &rd_in,
                         double &out) {
  TaskGeom t;
  task_geom(a, routine, cell_ord, part, t);
  bool blocked = false;
  Reader rd = rd_in;
  rd.comp = t.s.comp;
  rd.blocked = &blocked;
  double val, dv[3];
  value_derivs(rd, t.px, t.py, t.pz, t.ix, t.iy, t.iz, t.gx, t.gy, t.gz, t.s.comp, val, dv);
  if (blocked) return false;
  double surfv;
  if (a.tangent) {
    double g[BK_N];
    surface_term_grad(a, routine, t.s.comp, t.e, g);
    surfv = 0.0;
    for (int q = 0; q < BK_N; ++q) surfv += g[q] * a.kin_dot[q];
  } else {
    surfv = surface_term(a, routine, part, t.s.comp, t.e);
  }
  out = task_value(t, val, dv, surfv);
  if (a.tangent && a.addend) {
    const int nparts = (routine == R_P1 || routine == R_PF || routine == R_PG) ? 1 : 6;
    out += a.addend[cell_ord * nparts + part];
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

// the components a routine interpolates: p (0) or ut, vt, wt (1..3)
__host__ __device__ inline bool reads_comp(int routine, int comp) {
  const bool pressure = (routine == R_P1 || routine == R_PF || routine == R_PG);
  return pressure ? comp == 0 : comp != 0;
}

// Recording (the adjoint driver replays a step with it): a copy of the
// arrays the routine interpolates, taken before it runs, in the record of
// the current slot
static ForceRec *record_inputs(Block &b, int routine) {
  AdWork &w = b.ad;
  const size_t idx = (size_t)w.rec_slot * 6 + routine;
  if (w.recs.size() <= idx) w.recs.resize(idx + 1);
  ForceRec &r = w.recs[idx];
  const double *src[4] = {b.f[BF_P], b.f[BF_UT], b.f[BF_VT], b.f[BF_WT]};
  for (int c = 0; c < 4; ++c) {
    if (!reads_comp(routine, c)) continue;
    if (!r.arr[c] && dev_alloc(&r.arr[c], b.n2)) return nullptr;
    if (hipMemcpyAsync(r.arr[c], src[c], sizeof(double) * b.n2, hipMemcpyDeviceToDevice, rt().stream) !=
        hipSuccess) {
      set_error("forcing record: copy failed");
      return nullptr;
    }
  }
  r.live = false;
  return &r;
}

void ad_release_records(Block &b) {
  for (ForceRec &r : b.ad.recs) {
    for (auto &p : r.arr) if (p) (void)hipFree(p);
    if (r.tout) (void)hipFree(r.tout);
  }
  b.ad.recs.clear();
  b.ad.rec_slot = -1;
}

// The arguments of one forcing routine on block b (fields: the block's
// current field pointers).  Fails for a stale IB list; n_cells 0 = nothing
// to do.
static int forcing_args(Block *b, int routine, const double *kin, const double *kin_dot, FArgs &a, int &nparts,
                        int64_t &n_cells) {
  IbData &ib = b->ib;
  if (!ib.present) { set_error("block has no immersed body"); return -1; }
  const bool ts = (routine == R_VG || routine == R_PG);
  n_cells = ts ? b->counts[BC_TS] : b->counts[BC_IB];
  if (n_cells == 0) return 0;
  if (!ts && ib.n_mapped != n_cells) {
    set_error("forcing: IB list is stale (cell_count_solid after retagging is required)");
    return -1;
  }
  nparts = (routine == R_P1 || routine == R_PF || routine == R_PG) ? 1 : 6;
  a = FArgs{};
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
  a.tangent = kin_dot != nullptr;
  for (int q = 0; q < BK_N; ++q) a.kin_dot[q] = kin_dot ? kin_dot[q] : 0.0;
  a.ib_map = ib.ib_map;
  a.fld[0] = b->f[BF_P]; a.fld[1] = b->f[BF_UT]; a.fld[2] = b->f[BF_VT]; a.fld[3] = b->f[BF_WT];
  a.plain[0] = b->f[BF_P]; a.plain[1] = b->f[BF_U]; a.plain[2] = b->f[BF_V]; a.plain[3] = b->f[BF_W];
  const int64_t ntask = n_cells * nparts;
  if (grow(&ib.task_out, &ib.cap_task, ntask + 2)) return -1;
  if (grow(&ib.task_done, &ib.cap_done, ntask + 2)) return -1;
  a.miss = (int *)(ib.task_out + ib.cap_task - 1);  // 2 ints in the last slot
  BCFD_HIP(hipMemsetAsync(a.miss, 0, 2 * sizeof(int), rt().stream));
  return 0;
}

// The dependency waves of an in-place routine: task_done[t] = the wave that
// computed task t (1, 2, ...), task_out its value; `levels` the last wave.
static int force_waves(Block *b, const FArgs &a, int routine, int nparts, int &levels) {
  IbData &ib = b->ib;
  hipStream_t s = rt().stream;
  const int64_t ntask = a.n_cells * nparts;
  const unsigned g = div_up(ntask, 256);
  int *counters = a.miss;
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
  levels = pass;
  return 0;
}

static int force_misses(const FArgs &a) {
  int *h = (int *)rt().h_pinned_i;
  BCFD_HIP(hipMemcpyAsync(h, a.miss, sizeof(int), hipMemcpyDeviceToHost, rt().stream));
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  if (h[0] != 0) {
    set_error("forcing: %d image-point lookups fell outside their search window", h[0]);
    return -2;
  }
  return 0;
}

// Runs one forcing routine on block b.  kin_dot: null for the routine
// itself; otherwise the tangent-linear pass, with the block's field pointers
// set to the tangents by the caller.
static int force_block(Block *b, int routine, const double *kin, const double *kin_dot,
                       const double *addend = nullptr) {
  IbData &ib = b->ib;
  FArgs a;
  int nparts = 1;
  int64_t n_cells = 0;
  if (forcing_args(b, routine, kin, kin_dot, a, nparts, n_cells)) return -1;
  if (n_cells == 0) return 0;
  a.addend = kin_dot ? addend : nullptr;
  const bool ts = (routine == R_VG || routine == R_PG);
  hipStream_t s = rt().stream;
  const int64_t ntask = n_cells * nparts;
  const unsigned g = div_up(ntask, 256);
  ForceRec *rec = nullptr;
  if (!kin_dot && b->ad.rec_slot >= 0) {
    if (!(rec = record_inputs(*b, routine))) return -1;
  }
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
    int levels = 0;
    if (force_waves(b, a, routine, nparts, levels)) return -1;
    double *f0 = b->f[BF_P];
    double *f1 = b->f[BF_UT], *f2 = b->f[BF_VT], *f3 = b->f[BF_WT];
    if (rec) {
      if (grow(&rec->tout, &rec->cap_tout, ntask)) return -1;
      BCFD_HIP(hipMemcpyAsync(rec->tout, ib.task_out, sizeof(double) * ntask, hipMemcpyDeviceToDevice, s));
    }
    hipLaunchKernelGGL(k_force_scatter, dim3(g), dim3(256), 0, s, a, routine, nparts, ib.task_out, f0, f1, f2, f3);
    BCFD_CHECK_LAUNCH();
  }
  if (rec) {
    rec->ntask = ntask;
    rec->live = true;
  }
  return force_misses(a);
}

static int force(int id, int routine, const double *kin) {
  Block *b = get_block(id);
  if (!b) return -1;
  if (!b->ib.present) { set_error("block %d has no immersed body", id); return -1; }
  return force_block(b, routine, kin, nullptr);
}

// ---------------- stress_calculation (StressCalc.f90:14-208) ----------------
// the image point of element e and the cube of each component there
struct StressGeom {
  double diag;
  int cube[4][3];
};

__device__ inline void stress_geom(const FArgs &a, int64_t e, StressGeom &g) {
  const Dims &d = a.d;
  const double nrm[3] = {a.el[BE_COSA][e], a.el[BE_COSB][e], a.el[BE_COSG][e]};
  const double xc = a.el[BE_XCENT][e], yc = a.el[BE_YCENT][e], zc = a.el[BE_ZCENT][e];
  const int ic = find_index(xc, a.g[BG_X1], d.nx + 3, 2, d.nx + 1, a.miss);
  const int jc = find_index(yc, a.g[BG_Y1], d.ny + 3, 2, d.ny + 1, a.miss);
  const int kc = find_index(zc, a.g[BG_Z1], d.nz + 3, 2, d.nz + 1, a.miss);
  const double dX = a.g[BG_X1][ic] - a.g[BG_X1][ic - 1];
  const double dY = a.g[BG_Y1][jc] - a.g[BG_Y1][jc - 1];
  const double dZ = a.g[BG_Z1][kc] - a.g[BG_Z1][kc - 1];
  g.diag = sqrt(dX * dX + dY * dY + dZ * dZ);
  const double px = xc + g.diag * nrm[0], py = yc + g.diag * nrm[1], pz = zc + g.diag * nrm[2];
  for (int c = 0; c < 4; ++c) {
    const double *gx, *gy, *gz;
    int lx, ly, lz;
    comp_grids(a, c, gx, gy, gz, lx, ly, lz);
    g.cube[c][0] = find_index(px, gx, lx, 2, d.nx + 1, a.miss);
    g.cube[c][1] = find_index(py, gy, ly, 2, d.ny + 1, a.miss);
    g.cube[c][2] = find_index(pz, gz, lz, 2, d.nz + 1, a.miss);
  }
}

// The six terms of element e (viscous x, pressure x, viscous y, pressure y,
// projected areas xz and yz) from its centroid c, normal nrm, length alen,
// the kinematic vector and the fields the reader returns (rd.comp selects
// the component: 0 p, 1 u, 2 v, 3 w), at the forward's cubes; generic in the
// scalar type (Dual for the derivatives).
template <class T, class R>
__device__ inline void stress_terms_T(const FArgs &a, int e, const T c[3], const T nrm[3], T alen,
                                      const StressGeom &sg, const T k[BK_N], R &rd, double mu_f, double rho_f,
                                      T out[6]) {
  const T xc = c[0], yc = c[1], zc = c[2];
  const T px = xc + sg.diag * nrm[0], py = yc + sg.diag * nrm[1], pz = zc + sg.diag * nrm[2];
  const int sid = a.surf[e];
  T td = as<T>(0.0), tdd = as<T>(0.0);
  if (sid == 51) { td = k[BK_THETADOT1]; tdd = k[BK_THETADDOT1]; }
  else if (sid == 52) { td = k[BK_THETADOT2]; tdd = k[BK_THETADDOT2]; }
  const T surf[3] = {k[BK_XDOT], -td * (zc - k[BK_PIVZ]) + k[BK_YDOT], td * (yc - k[BK_PIVY])};
  const T ac_z = -(td * td) * (zc - k[BK_PIVZ]), ac_y = -(td * td) * (yc - k[BK_PIVY]);
  const T at_z = tdd * (yc - k[BK_PIVY]), at_y = -tdd * (zc - k[BK_PIVZ]);
  const T dpdn = -((ac_z + at_z) * nrm[2] + (ac_y + at_y) * nrm[1]) - k[BK_YDDOT] * nrm[1];
  T ddn[3];
  for (int q = 0; q < 3; ++q) {
    const double *gx, *gy, *gz;
    int lx, ly, lz;
    comp_grids(a, q + 1, gx, gy, gz, lx, ly, lz);
    rd.comp = q + 1;
    T val, dv[3];
    value_derivs<R, T>(rd, px, py, pz, sg.cube[q + 1][0], sg.cube[q + 1][1], sg.cube[q + 1][2], gx, gy, gz, q + 1,
                       val, dv);
    const T dn = dv[0] * nrm[0] + dv[1] * nrm[1] + dv[2] * nrm[2];
    ddn[q] = (2.0 / sg.diag) * (val - surf[q]) - dn;
  }
  const T area = alen / 2.0;
  const T area_yz = 0.5 * fabs(alen * nrm[0]);
  const T area_xz = 0.5 * fabs(alen * nrm[1]);
  const T dd = ddn[0] * nrm[0] + ddn[1] * nrm[1] + ddn[2] * nrm[2];
  const T sh0 = (ddn[0] - dd * nrm[0]) * mu_f * area;
  const T sh1 = (ddn[1] - dd * nrm[1]) * mu_f * area;
  const double *gx, *gy, *gz;
  int lx, ly, lz;
  comp_grids(a, 0, gx, gy, gz, lx, ly, lz);
  rd.comp = 0;
  T val, dv[3];
  value_derivs<R, T>(rd, px, py, pz, sg.cube[0][0], sg.cube[0][1], sg.cube[0][2], gx, gy, gz, 0, val, dv);
  const T dpdn_e = dv[0] * nrm[0] + dv[1] * nrm[1] + dv[2] * nrm[2];
  const T cval = val - (dpdn_e + dpdn) * sg.diag * 0.5;
  out[0] = sh0;
  out[1] = -cval * area * nrm[0] * rho_f;
  out[2] = sh1;
  out[3] = -cval * area * nrm[1] * rho_f;
  out[4] = area_xz;
  out[5] = area_yz;
}

__global__ void k_stress_terms(FArgs a, double mu_f, double rho_f, double *terms) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.n_cells) return;
  StressGeom sg;
  stress_geom(a, e, sg);
  double c[3], nrm[3];
  c[0] = a.el[BE_XCENT][e]; c[1] = a.el[BE_YCENT][e]; c[2] = a.el[BE_ZCENT][e];
  nrm[0] = a.el[BE_COSA][e]; nrm[1] = a.el[BE_COSB][e]; nrm[2] = a.el[BE_COSG][e];
  Reader rd{&a, 0, false, 0, 0, nullptr, nullptr, 0, nullptr};
  stress_terms_T<double>(a, (int)e, c, nrm, a.el[BE_ELEN][e], sg, a.kin, rd, mu_f, rho_f, terms + 6 * e);
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

// ---------------------------------------------------------------------------
// Tangent-linear and adjoint passes of the forcing routines and of the
// body-force sums (the adjoint time step; no counterpart in the reference).
// Geometry frozen: IB cells, nearest elements, distances, normals and
// image-point cubes are the forward's.  A reconstruction is then linear in
// the eight values of its cube and in its surface term, and the surface
// term is a function of the kinematic vector (surface_term_grad).
// ---------------------------------------------------------------------------

// 1 at one position, 0 elsewhere: value_derivs through it gives the
// image-point value and derivatives per unit value at that cube corner
struct UnitReader {
  int ti, tj, tk;
  __device__ double operator()(int ii, int jj, int kk) const {
    return (ii == ti && jj == tj && kk == tk) ? 1.0 : 0.0;
  }
};

// corner m (bit 0: i + 1, bit 1: j + 1, bit 2: k + 1) of the cube
// value_derivs reads for component comp at cube index (i, j, k)
__device__ inline void cube_corner(int comp, int i, int j, int k, int m, int &ii, int &jj, int &kk) {
  ii = i - (comp == 1) + (m & 1);
  jj = j - (comp == 2) + ((m >> 1) & 1);
  kk = k - (comp == 3) + ((m >> 2) & 1);
}

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// adds c[q] summed over the wavefront into acc[q] (one atomic per entry
// and wavefront; every lane of the wavefront calls it)
__device__ inline void wave_add_kin(const double c[BK_N], double *acc) {
  const int lane = threadIdx.x & 63;
  for (int q = 0; q < BK_N; ++q) {
    const double v = wave_sum(c[q]);
    if (lane == 0 && v != 0.0) atomicAdd(&acc[q], v);
  }
}

// location written by task t and whether t is its last writer in serial
// order (the opposite face of the neighbouring IB cell writes the same
// location for the velocity routines, as k_force_scatter decides)
__device__ inline int64_t task_target(const FArgs &a, int routine, int nparts, int64_t t, int &comp, bool &last) {
  const int64_t n = t / nparts;
  const int part = (int)(t % nparts);
  const PartSpec s = part_spec(routine, part);
  const int i = a.ijk[n] + s.tx, j = a.ijk[n + a.cap] + s.ty, k = a.ijk[n + 2 * a.cap] + s.tz;
  const Dims &d = a.d;
  comp = s.comp;
  last = true;
  if (routine == R_V1 || routine == R_VF) {
    const int si = (s.comp == 1), sj = (s.comp == 2), sk = (s.comp == 3);
    long long other = -1;
    if ((part % 2) == 0) {
      const int ni = i + si, nj = j + sj, nk = k + sk;
      if (ni <= d.NX && nj <= d.NY && nk <= d.NZ) {
        const int m = a.ib_map[ix3(d, ni, nj, nk)];
        if (m >= 0) other = 6LL * m + part + 1;
      }
    } else {
      const int m = a.ib_map[ix3(d, i, j, k)];
      if (m >= 0) other = 6LL * m + part - 1;
    }
    last = !(other > t);
  }
  return ix3(d, i, j, k);
}

// ybar[t] = adjoint of the written location for its last writer, 0 for the
// others (their values were overwritten)
struct Lam4 {
  double *f[4];   // p, u, v, w components
};

__global__ void k_force_adj_init(FArgs a, int routine, int nparts, double *ybar, Lam4 lam_out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_cells * nparts) return;
  int comp;
  bool last;
  const int64_t loc = task_target(a, routine, nparts, t, comp, last);
  ybar[t] = last ? lam_out.f[comp][loc] : 0.0;
}

__global__ void k_force_adj_clear(FArgs a, int routine, int nparts, Lam4 lam_out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_cells * nparts) return;
  int comp;
  bool last;
  const int64_t loc = task_target(a, routine, nparts, t, comp, last);
  lam_out.f[comp][loc] = 0.0;
}

// The values a recorded forcing call read: the arrays at its start, and for
// the in-place routines the earlier tasks' outputs where they wrote first
// (the forward's reads, Reader).
struct RecView {
  const double *arr[4];
  const double *tout;
  bool live;
};

struct RecReader {
  const FArgs *a;
  const RecView *rv;
  Reader w;   // writer lookup of the in-place routines (ordered) or none
  __device__ double operator()(int ii, int jj, int kk) const {
    const int64_t loc = ix3(a->d, ii, jj, kk);
    if (w.ordered) {
      const long long wr = w.writer_before(loc, ii, jj, kk);
      if (wr >= 0) return rv->tout[wr];
    }
    return rv->arr[w.comp][loc];
  }
};

__device__ inline void elem_geom(const FArgs &a, int e, double c[3], double nrm[3]) {
  c[0] = a.el[BE_XCENT][e]; c[1] = a.el[BE_YCENT][e]; c[2] = a.el[BE_ZCENT][e];
  nrm[0] = a.el[BE_COSA][e]; nrm[1] = a.el[BE_COSB][e]; nrm[2] = a.el[BE_COSG][e];
}

// d(task output) / d(centroid, normal of its element) . (dc, dn) with the
// recorded values: one Dual evaluation of task_eval
__device__ inline double task_geom_derivative(const FArgs &a, int routine, int part, int64_t cell_ord,
                                              const TaskGeom &g, const RecView &rv, long long key,
                                              const double dg[6]) {
  double c0[3], n0[3];
  elem_geom(a, g.e, c0, n0);
  Dual c[3], nrm[3], kin[BK_N];
  for (int q = 0; q < 3; ++q) {
    c[q] = {c0[q], dg[q]};
    nrm[q] = {n0[q], dg[3 + q]};
  }
  for (int q = 0; q < BK_N; ++q) kin[q] = {a.kin[q], 0.0};
  const bool ordered = (routine == R_V1 || routine == R_P1 || routine == R_PF);
  const RecReader rd{&a, &rv, Reader{&a, g.s.comp, ordered, routine, key, nullptr, nullptr, 0, nullptr}};
  return task_eval<Dual>(a, routine, part, cell_ord, g, c, nrm, kin, rd).t;
}

// The tasks of one dependency wave (level; 0: all tasks, for the routines
// without in-place reads), last wave first: each passes ybar[t] times its
// cube weights to the earlier task that wrote a cube position (in-place
// routines) or to the adjoint of the array read, and times its surface
// weight through the surface term's gradient into kin_bar; with a record
// (rv.live), ybar[t] times d(output)/d(centroid, normal) into the element's
// geometric adjoint gbar (7 x ne).
__global__ void k_force_adj(FArgs a, int routine, int nparts, int level, const int *task_done, double *ybar,
                            Lam4 lam_in, double *kin_bar, RecView rv, double *gbar, int64_t ne) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double kc[BK_N];
  for (int q = 0; q < BK_N; ++q) kc[q] = 0.0;
  const bool active = t < a.n_cells * nparts && (level == 0 || task_done[t] == level);
  const double yb = active ? ybar[t] : 0.0;
  if (active && yb != 0.0) {
    TaskGeom g;
    const int64_t n = t / nparts;
    const int part = (int)(t % nparts);
    task_geom(a, routine, n, part, g);
    const double s2 = g.sur2, n1 = g.n1;
    double dval, ddn, dsurf;
    if (g.s.comp == 0) {
      dval = 1.0;
      ddn = s2 * s2 / (2.0 * n1) - 0.5 * n1;
      dsurf = -(s2 * s2) / (2.0 * n1) + s2 - 0.5 * n1;
    } else {
      dval = -(s2 * s2) / (n1 * n1) + 2.0 * s2 / n1;
      ddn = s2 * s2 / n1 - s2;
      dsurf = s2 * s2 / (n1 * n1) - 2.0 * s2 / n1 + 1.0;
    }
    if (g.ghost) ddn = -ddn;
    const bool ordered = (routine == R_V1 || routine == R_P1 || routine == R_PF);
    const Reader rd{&a, g.s.comp, true, routine, (long long)t, nullptr, nullptr, 0, nullptr};
    for (int m = 0; m < 8; ++m) {
      int ii, jj, kk;
      cube_corner(g.s.comp, g.ix, g.iy, g.iz, m, ii, jj, kk);
      const UnitReader ur{ii, jj, kk};
      double val, dv[3];
      value_derivs(ur, g.px, g.py, g.pz, g.ix, g.iy, g.iz, g.gx, g.gy, g.gz, g.s.comp, val, dv);
      const double w = dval * val + ddn * (dv[0] * g.ca + dv[1] * g.cb + dv[2] * g.cg);
      if (w == 0.0) continue;
      const int64_t loc = ix3(a.d, ii, jj, kk);
      const long long wr = ordered ? rd.writer_before(loc, ii, jj, kk) : -1;
      if (wr >= 0) atomicAdd(&ybar[wr], w * yb);
      else atomicAdd(&lam_in.f[g.s.comp][loc], w * yb);
    }
    double gs[BK_N];
    surface_term_grad(a, routine, g.s.comp, g.e, gs);
    for (int q = 0; q < BK_N; ++q) kc[q] = dsurf * yb * gs[q];
    if (rv.live) {
      for (int q = 0; q < 6; ++q) {
        double dg[6] = {0, 0, 0, 0, 0, 0};
        dg[q] = 1.0;
        const double d = task_geom_derivative(a, routine, part, n, g, rv, (long long)t, dg);
        if (d != 0.0) atomicAdd(&gbar[q * ne + g.e], yb * d);
      }
    }
  }
  wave_add_kin(kc, kin_bar);
}

// the geometric tangent of every task: d(output)/d(centroid, normal) . the
// element tangents (gdot, 7 x ne), with the recorded values
__global__ void k_force_geom_tlm(FArgs a, int routine, int nparts, RecView rv, const double *gdot, int64_t ne,
                                 double *addend) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_cells * nparts) return;
  TaskGeom g;
  const int64_t n = t / nparts;
  const int part = (int)(t % nparts);
  task_geom(a, routine, n, part, g);
  double dg[6];
  for (int q = 0; q < 6; ++q) dg[q] = gdot[q * ne + g.e];
  addend[t] = task_geom_derivative(a, routine, part, n, g, rv, (long long)t, dg);
}

// the ghost routines' adjoint seeds: ybar[t] = adjoint of the reconstructed
// ghost; the adjoint of the copied ("t") ghost goes to the copied field
// entry; both ghost adjoints are consumed
struct Gh14 {
  double *g[BGH_N];
};

__global__ void k_force_adj_ghost(FArgs a, int routine, int nparts, double *ybar, Lam4 lam_plain, Gh14 gh) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n_cells * nparts) return;
  const int64_t n = t / nparts;
  const int part = (int)(t % nparts);
  const PartSpec s = part_spec(routine, part);
  const int i = a.ijk[n] + s.tx, j = a.ijk[n + a.cap] + s.ty, k = a.ijk[n + 2 * a.cap] + s.tz;
  ybar[t] = gh.g[s.ghost][n];
  gh.g[s.ghost][n] = 0.0;
  const double c = gh.g[s.ghost_t][n];
  gh.g[s.ghost_t][n] = 0.0;
  if (c != 0.0) atomicAdd(&lam_plain.f[s.comp][ix3(a.d, i, j, k)], c);
}

static int rec_view(Block &b, int routine, int rec, RecView &rv) {
  rv = RecView{{nullptr, nullptr, nullptr, nullptr}, nullptr, false};
  if (rec < 0) return 0;
  const size_t idx = (size_t)rec * 6 + routine;
  if (idx >= b.ad.recs.size() || !b.ad.recs[idx].live) {
    set_error("forcing: no record of routine %d in slot %d", routine, rec);
    return -1;
  }
  const ForceRec &r = b.ad.recs[idx];
  for (int c = 0; c < 4; ++c) rv.arr[c] = r.arr[c];
  rv.tout = r.tout;
  rv.live = true;
  return 0;
}

int forcing_tlm(Block &b, int routine, const double kin[BK_N], const double kin_dot[BK_N], int rec) {
  const double *addend = nullptr;
  if (rec >= 0) {
    // the geometric tangents, from the recorded values and the element tangents
    RecView rv;
    if (rec_view(b, routine, rec, rv) || ad_geometry(b)) return -1;
    if (!b.ad.gdot_set) { set_error("forcing_tlm: no geometric tangent set"); return -1; }
    FArgs a;
    int nparts = 1;
    int64_t n_cells = 0;
    if (forcing_args(&b, routine, kin, nullptr, a, nparts, n_cells)) return -1;
    if (n_cells > 0) {
      const int64_t ntask = n_cells * nparts;
      if (grow(&b.ad.addend, &b.ad.cap_addend, ntask)) return -1;
      hipLaunchKernelGGL(k_force_geom_tlm, dim3(div_up(ntask, 256)), dim3(256), 0, rt().stream, a, routine, nparts,
                         rv, b.ad.gdot, b.ib.n_elems, b.ad.addend);
      BCFD_CHECK_LAUNCH();
      if (force_misses(a)) return -2;
      addend = b.ad.addend;
    }
  }
  SwapFields sw(b);
  if (routine != R_VG && routine != R_PG) return force_block(&b, routine, kin, kin_dot, addend);
  // the ghost routines write the ghost arrays: their tangents
  if (ad_ghosts(b)) return -1;
  for (int q = 0; q < BGH_N; ++q) std::swap(b.ib.ghost[q], b.ad.gh[q]);
  const int rc = force_block(&b, routine, kin, kin_dot, addend);
  for (int q = 0; q < BGH_N; ++q) std::swap(b.ib.ghost[q], b.ad.gh[q]);
  return rc;
}

int forcing_adj(Block &b, int routine, const double kin[BK_N], int rec) {
  RecView rv;
  if (rec_view(b, routine, rec, rv)) return -1;
  if (rec >= 0 && ad_geometry(b)) return -1;
  FArgs a;
  int nparts = 1;
  int64_t n_cells = 0;
  if (forcing_args(&b, routine, kin, nullptr, a, nparts, n_cells)) return -1;
  if (n_cells == 0) return 0;
  IbData &ib = b.ib;
  hipStream_t s = rt().stream;
  const int64_t ntask = n_cells * nparts;
  const unsigned g = div_up(ntask, 256);
  int levels = 0;
  const bool ordered = (routine == R_V1 || routine == R_P1 || routine == R_PF);
  if (ordered && force_waves(&b, a, routine, nparts, levels)) return -1;
  if (grow(&b.ad.ybar, &b.ad.cap_ybar, ntask)) return -1;
  // adjoints of the written arrays and of the arrays the cubes read
  Lam4 out{{b.ad.s[BF_P], b.ad.s[BF_UT], b.ad.s[BF_VT], b.ad.s[BF_WT]}};
  Lam4 in = out;
  if (routine == R_VF) { out.f[1] = b.ad.s[BF_U]; out.f[2] = b.ad.s[BF_V]; out.f[3] = b.ad.s[BF_W]; }
  if (routine == R_VG || routine == R_PG) {
    if (ad_ghosts(b)) return -1;
    Gh14 gh;
    for (int q = 0; q < BGH_N; ++q) gh.g[q] = b.ad.gh[q];
    const Lam4 plain{{b.ad.s[BF_P], b.ad.s[BF_U], b.ad.s[BF_V], b.ad.s[BF_W]}};
    hipLaunchKernelGGL(k_force_adj_ghost, dim3(g), dim3(256), 0, s, a, routine, nparts, b.ad.ybar, plain, gh);
  } else {
    hipLaunchKernelGGL(k_force_adj_init, dim3(g), dim3(256), 0, s, a, routine, nparts, b.ad.ybar, out);
    hipLaunchKernelGGL(k_force_adj_clear, dim3(g), dim3(256), 0, s, a, routine, nparts, out);
  }
  const int64_t ne = ib.n_elems;
  if (!ordered) {
    hipLaunchKernelGGL(k_force_adj, dim3(g), dim3(256), 0, s, a, routine, nparts, 0, ib.task_done, b.ad.ybar, in,
                       b.ad.kin, rv, b.ad.gbar, ne);
  } else {
    for (int level = levels; level >= 1; --level)
      hipLaunchKernelGGL(k_force_adj, dim3(g), dim3(256), 0, s, a, routine, nparts, level, ib.task_done, b.ad.ybar,
                         in, b.ad.kin, rv, b.ad.gbar, ne);
  }
  BCFD_CHECK_LAUNCH();
  return force_misses(a);
}

// ---- body-force sums: tangent-linear and adjoint ----
// S = (viscous x, pressure x, viscous y, pressure y) summed over elements,
// the quantities stress_calculation divides by its projected areas; the
// terms are stress_terms_T on Dual values: fields (value, tangent), the
// kinematic vector (value, kin_dot) and, with geometry, the element's
// centroid, normal and length (value, tangent).

// field values with their tangents (tangent-linear) or with a unit seed at
// one cube position (adjoint)
struct DualFieldReader {
  const FArgs *a;
  const double *tan[4];   // null: no tangents
  int comp;
  int scomp;               // seeded component (-1: none)
  int64_t sloc;
  __device__ Dual operator()(int ii, int jj, int kk) const {
    const int64_t loc = ix3(a->d, ii, jj, kk);
    const double v = a->fld[comp][loc];
    if (tan[0]) return {v, tan[comp][loc]};
    return {v, (comp == scomp && loc == sloc) ? 1.0 : 0.0};
  }
};

__device__ inline void stress_dual_inputs(const FArgs &a, int64_t e, const double *gdot, int64_t ne, int gseed,
                                          const double *kdot, int kseed, Dual c[3], Dual nrm[3], Dual &alen,
                                          Dual kin[BK_N]) {
  const double cv[3] = {a.el[BE_XCENT][e], a.el[BE_YCENT][e], a.el[BE_ZCENT][e]};
  const double nv[3] = {a.el[BE_COSA][e], a.el[BE_COSB][e], a.el[BE_COSG][e]};
  for (int q = 0; q < 3; ++q) {
    c[q] = {cv[q], gdot ? gdot[q * ne + e] : (gseed == q ? 1.0 : 0.0)};
    nrm[q] = {nv[q], gdot ? gdot[(3 + q) * ne + e] : (gseed == 3 + q ? 1.0 : 0.0)};
  }
  alen = {a.el[BE_ELEN][e], gdot ? gdot[6 * ne + e] : (gseed == 6 ? 1.0 : 0.0)};
  for (int q = 0; q < BK_N; ++q) kin[q] = {a.kin[q], kdot ? kdot[q] : (kseed == q ? 1.0 : 0.0)};
}

// tangents of the four terms: fields (tangent arrays), kin_dot and, if gdot,
// the geometric tangents
__global__ void k_stress_tlm_terms(FArgs a, double mu_f, double rho_f, Lam4 tan, const double *gdot, int64_t ne,
                                   double *terms) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.n_cells) return;
  StressGeom sg;
  stress_geom(a, e, sg);
  Dual c[3], nrm[3], alen, kin[BK_N];
  stress_dual_inputs(a, e, gdot, ne, -1, a.kin_dot, -1, c, nrm, alen, kin);
  if (!gdot) {
    for (int q = 0; q < 3; ++q) { c[q].t = 0.0; nrm[q].t = 0.0; }
    alen.t = 0.0;
  }
  DualFieldReader rd{&a, {tan.f[0], tan.f[1], tan.f[2], tan.f[3]}, 0, -1, 0};
  Dual out[6];
  stress_terms_T<Dual>(a, (int)e, c, nrm, alen, sg, kin, rd, mu_f, rho_f, out);
  double *t = terms + 4 * e;
  for (int q = 0; q < 4; ++q) t[q] = out[q].t;
}

struct Sum4 {
  double v[4];
};

// sbar (the adjoints of the four sums) into the adjoints of u, v, w, p at
// the image-point cubes, into kin_bar and, with geometry, into gbar: one
// seeded Dual evaluation per cube position, kinematic entry and geometric
// input
__global__ void k_stress_adj(FArgs a, double mu_f, double rho_f, Sum4 sbar, Lam4 lam, double *kin_bar, double *gbar,
                             int64_t ne) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double kc[BK_N];
  for (int q = 0; q < BK_N; ++q) kc[q] = 0.0;
  if (e < a.n_cells) {
    StressGeom sg;
    stress_geom(a, e, sg);
    auto adj_of = [&](int gseed, int kseed, int scomp, int64_t sloc) {
      Dual c[3], nrm[3], alen, kin[BK_N];
      stress_dual_inputs(a, e, nullptr, ne, gseed, nullptr, kseed, c, nrm, alen, kin);
      DualFieldReader rd{&a, {nullptr, nullptr, nullptr, nullptr}, 0, scomp, sloc};
      Dual out[6];
      stress_terms_T<Dual>(a, (int)e, c, nrm, alen, sg, kin, rd, mu_f, rho_f, out);
      return sbar.v[0] * out[0].t + sbar.v[1] * out[1].t + sbar.v[2] * out[2].t + sbar.v[3] * out[3].t;
    };
    for (int comp = 0; comp < 4; ++comp)
      for (int m = 0; m < 8; ++m) {
        int ii, jj, kk;
        cube_corner(comp, sg.cube[comp][0], sg.cube[comp][1], sg.cube[comp][2], m, ii, jj, kk);
        const int64_t loc = ix3(a.d, ii, jj, kk);
        const double g = adj_of(-1, -1, comp, loc);
        if (g != 0.0) atomicAdd(&lam.f[comp][loc], g);
      }
    for (int q = 0; q < BK_N; ++q) kc[q] = adj_of(-1, q, -1, -1);
    if (gbar)
      for (int q = 0; q < 7; ++q) {
        const double g = adj_of(q, -1, -1, -1);
        if (g != 0.0) atomicAdd(&gbar[q * ne + e], g);
      }
  }
  wave_add_kin(kc, kin_bar);
}

// four sums over the elements (one workgroup, a fixed order)
__global__ void __launch_bounds__(256) k_sum4(const double *__restrict__ terms, int64_t ne, int stride,
                                              double *__restrict__ out) {
  __shared__ double part[4][256];
  const int t = threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t e = t; e < ne; e += 256)
    for (int q = 0; q < 4; ++q) acc[q] += terms[stride * e + q];
  for (int q = 0; q < 4; ++q) part[q][t] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
      for (int q = 0; q < 4; ++q) part[q][t] += part[q][t + w];
    __syncthreads();
  }
  if (t < 4) out[t] = part[t][0];
}

static int stress_args(Block &b, const double kin[BK_N], const double *kin_dot, FArgs &a) {
  IbData &ib = b.ib;
  if (!ib.present) { set_error("block has no immersed body"); return -1; }
  a = FArgs{};
  a.d = b.d;
  for (int q = 0; q < BG_NGRID; ++q) a.g[q] = b.g[q];
  a.n_cells = ib.n_elems;
  for (int q = 0; q < BE_NELEM; ++q) a.el[q] = ib.elem[q];
  a.surf = ib.surf;
  for (int q = 0; q < BK_N; ++q) { a.kin[q] = kin[q]; a.kin_dot[q] = kin_dot ? kin_dot[q] : 0.0; }
  a.tangent = kin_dot != nullptr;
  a.fld[0] = b.f[BF_P]; a.fld[1] = b.f[BF_U]; a.fld[2] = b.f[BF_V]; a.fld[3] = b.f[BF_W];
  if (grow(&ib.stress_work, &ib.cap_stress, 6 * ib.n_elems + 8)) return -1;
  a.miss = (int *)(ib.stress_work + 6 * ib.n_elems + 4);
  hipLaunchKernelGGL(k_stress_clear, dim3(1), dim3(64), 0, rt().stream, a.miss);
  return 0;
}

static int stress_misses(const FArgs &a) {
  int *h = (int *)rt().h_pinned_i;
  BCFD_HIP(hipMemcpyAsync(h, a.miss, sizeof(int), hipMemcpyDeviceToHost, rt().stream));
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  if (h[0] != 0) { set_error("stress: %d probe lookups missed", h[0]); return -2; }
  return 0;
}

int stress_tlm(Block &b, double mu_f, double rho_f, const double kin[BK_N], const double kin_dot[BK_N], int geom,
               double out[4]) {
  FArgs a;
  if (stress_args(b, kin, kin_dot, a)) return -1;
  if (geom) {
    if (ad_geometry(b)) return -1;
    if (!b.ad.gdot_set) { set_error("stress_tlm: no geometric tangent set"); return -1; }
  }
  IbData &ib = b.ib;
  double *terms = ib.stress_work, *res = ib.stress_work + 6 * ib.n_elems;
  hipStream_t s = rt().stream;
  const Lam4 tan{{b.ad.s[BF_P], b.ad.s[BF_U], b.ad.s[BF_V], b.ad.s[BF_W]}};
  hipLaunchKernelGGL(k_stress_tlm_terms, dim3(div_up(ib.n_elems, 256)), dim3(256), 0, s, a, mu_f, rho_f, tan,
                     geom ? b.ad.gdot : nullptr, ib.n_elems, terms);
  hipLaunchKernelGGL(k_sum4, dim3(1), dim3(256), 0, s, terms, ib.n_elems, 4, res);
  BCFD_CHECK_LAUNCH();
  BCFD_HIP(hipMemcpyAsync(rt().h_pinned, res, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
  if (stress_misses(a)) return -2;
  for (int q = 0; q < 4; ++q) out[q] = rt().h_pinned[q];
  return 0;
}

int stress_adj(Block &b, double mu_f, double rho_f, const double kin[BK_N], const double sbar[4], int geom) {
  FArgs a;
  if (stress_args(b, kin, nullptr, a)) return -1;
  if (geom && ad_geometry(b)) return -1;
  const Sum4 sb{{sbar[0], sbar[1], sbar[2], sbar[3]}};
  const Lam4 lam{{b.ad.s[BF_P], b.ad.s[BF_U], b.ad.s[BF_V], b.ad.s[BF_W]}};
  hipLaunchKernelGGL(k_stress_adj, dim3(div_up(b.ib.n_elems, 256)), dim3(256), 0, rt().stream, a, mu_f, rho_f, sb,
                     lam, b.ad.kin, geom ? b.ad.gbar : nullptr, b.ib.n_elems);
  BCFD_CHECK_LAUNCH();
  return stress_misses(a);
}

// the four sums of the forward (a fixed order; the adjoint workload's
// objective reads their x parts)
int stress_sums(Block &b, double mu_f, double rho_f, const double kin[BK_N], double out[4]) {
  FArgs a;
  if (stress_args(b, kin, nullptr, a)) return -1;
  IbData &ib = b.ib;
  double *terms = ib.stress_work, *res = ib.stress_work + 6 * ib.n_elems;
  hipStream_t s = rt().stream;
  hipLaunchKernelGGL(k_stress_terms, dim3(div_up(ib.n_elems, 256)), dim3(256), 0, s, a, mu_f, rho_f, terms);
  hipLaunchKernelGGL(k_sum4, dim3(1), dim3(256), 0, s, terms, ib.n_elems, 6, res);
  BCFD_CHECK_LAUNCH();
  BCFD_HIP(hipMemcpyAsync(rt().h_pinned, res, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
  if (stress_misses(a)) return -2;
  for (int q = 0; q < 4; ++q) out[q] = rt().h_pinned[q];
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
