// Exact nearest-element search through a uniform grid of the element
// centroids, for the immersed-boundary searches of Search.F90
// (selectiveRetagging_th :520-651 via tagging_th_core :305-338,
// computeNormDistance :762-858).
//
// The reference scans every element m = 1..E and keeps the first minimum of
// the squared distance (strict <, starting from 1e14); that is the
// lexicographic minimum of (distance, index) when it is below 1e14, else 0.
// Here one wavefront serves one query cell: the centroids sit in a uniform
// grid of bins (counting-sorted by bin), the lanes scan the bins of growing
// Chebyshev shells r = 0, 1, 2, ... around the bin of the query's first point
// and keep lexicographic minima, which the wavefront reduces after each
// shell.  A point q whose bin is dl shells from the centre bin has every
// element outside the scanned cube at least (r - dl) h away (h = bin side;
// a point outside the grid is clamped into it and the bound still holds),
// so once each point's best squared distance is below ((r - dl) h)^2, with
// a relative margin of 1e-9 for rounding, no element outside can be closer
// or equally close, and the result equals the reference's scan bit for bit
// (same distance expression, same tie-break).
//
// Host builds of this header run the same lane bodies in tests.
#pragma once

#include <cmath>
#include <cstdint>

namespace bcfd {

constexpr int kBinDiv = 32;                 // bins along the longest extent of the centroids
constexpr int kBinAxis = kBinDiv + 1;       // bins per axis at most
constexpr int kMaxBins = kBinAxis * kBinAxis * kBinAxis;
constexpr int kSearchLanes = 64;            // lanes per query (one wavefront)

struct BinGrid {
  double ox, oy, oz, h;  // origin (minimum corner of the centroids) and bin side
  int nx, ny, nz;
};

// the bin grid of centroids with bounding box [mn, mx]
__host__ __device__ inline BinGrid bin_grid_of(const double mn[3], const double mx[3]) {
  BinGrid g;
  g.ox = mn[0];
  g.oy = mn[1];
  g.oz = mn[2];
  double ext = mx[0] - mn[0];
  ext = mx[1] - mn[1] > ext ? mx[1] - mn[1] : ext;
  ext = mx[2] - mn[2] > ext ? mx[2] - mn[2] : ext;
  g.h = ext > 0.0 ? ext / kBinDiv : 1.0;
  int n[3];
  for (int a = 0; a < 3; ++a) {
    const double t = std::floor((mx[a] - mn[a]) / g.h) + 1.0;
    n[a] = t > kBinAxis ? kBinAxis : (t < 1.0 ? 1 : (int)t);
  }
  g.nx = n[0];
  g.ny = n[1];
  g.nz = n[2];
  return g;
}

// bin coordinate of x on one axis, clamped into the grid
__host__ __device__ inline int bin_axis(double x, double o, double h, int n) {
  double t = std::floor((x - o) / h);
  t = t < 0.0 ? 0.0 : (t > (double)(n - 1) ? (double)(n - 1) : t);
  return (int)t;
}

__host__ __device__ inline int bin_of(const BinGrid &g, double x, double y, double z) {
  return bin_axis(x, g.ox, g.h, g.nx) + g.nx * (bin_axis(y, g.oy, g.h, g.ny) + g.ny * bin_axis(z, g.oz, g.h, g.nz));
}

__host__ __device__ inline int iabs_(int v) { return v < 0 ? -v : v; }
__host__ __device__ inline int imax3(int a, int b, int c) { return a > b ? (a > c ? a : c) : (b > c ? b : c); }

// T query points of one cell; point 0 sets the centre bin
template <int T>
struct NearestQuery {
  double x[T], y[T], z[T];
  int cx, cy, cz;  // centre bin
  int dl[T];       // Chebyshev bin distance of point s from the centre bin
};

template <int T>
__host__ __device__ inline NearestQuery<T> make_query(const BinGrid &g, const double *x, const double *y,
                                                      const double *z) {
  NearestQuery<T> Q;
  for (int s = 0; s < T; ++s) {
    Q.x[s] = x[s];
    Q.y[s] = y[s];
    Q.z[s] = z[s];
  }
  Q.cx = bin_axis(x[0], g.ox, g.h, g.nx);
  Q.cy = bin_axis(y[0], g.oy, g.h, g.ny);
  Q.cz = bin_axis(z[0], g.oz, g.h, g.nz);
  for (int s = 0; s < T; ++s)
    Q.dl[s] = imax3(iabs_(bin_axis(x[s], g.ox, g.h, g.nx) - Q.cx), iabs_(bin_axis(y[s], g.oy, g.h, g.ny) - Q.cy),
                    iabs_(bin_axis(z[s], g.oz, g.h, g.nz) - Q.cz));
  return Q;
}

// the start value of a lane's minima: no element yet
constexpr double kNoDist = 1e300;
constexpr int kNoElem = 0x7fffffff;

__host__ __device__ inline bool lex_less(double d, int e, double bd, int be) { return d < bd || (d == bd && e < be); }

// one lane's part of shell r: the shell's bins c = lane, lane + nlanes, ...
// of the (2r+1)^3 cube, each element of each bin against every point
template <int T>
__host__ __device__ inline void shell_scan(const BinGrid &g, const int *bstart, const int *belem, const double *xc,
                                           const double *yc, const double *zc, const NearestQuery<T> &Q, int r,
                                           int lane, int nlanes, double *bd, int *be) {
  const int side = 2 * r + 1, ncube = side * side * side;
  for (int c = lane; c < ncube; c += nlanes) {
    const int dx = c % side - r, dy = (c / side) % side - r, dz = c / (side * side) - r;
    if (imax3(iabs_(dx), iabs_(dy), iabs_(dz)) != r) continue;
    const int x = Q.cx + dx, y = Q.cy + dy, z = Q.cz + dz;
    if (x < 0 || x >= g.nx || y < 0 || y >= g.ny || z < 0 || z >= g.nz) continue;
    const int b = x + g.nx * (y + g.ny * z);
    const int p1 = bstart[b + 1];
    for (int p = bstart[b]; p < p1; ++p) {
      const int e = belem[p];
      const double ex = xc[e], ey = yc[e], ez = zc[e];
      for (int s = 0; s < T; ++s) {
        // the reference's expression: dx^2 + dy^2 + dz^2, left to right
        const double ax = Q.x[s] - ex, ay = Q.y[s] - ey, az = Q.z[s] - ez;
        const double d = ax * ax + ay * ay + az * az;
        if (lex_less(d, e, bd[s], be[s])) {
          bd[s] = d;
          be[s] = e;
        }
      }
    }
  }
}

// after shell r (minima reduced over the lanes): are all points final?
template <int T>
__host__ __device__ inline bool shell_done(const BinGrid &g, const NearestQuery<T> &Q, int r, const double *bd) {
  if (Q.cx - r <= 0 && Q.cx + r >= g.nx - 1 && Q.cy - r <= 0 && Q.cy + r >= g.ny - 1 && Q.cz - r <= 0 &&
      Q.cz + r >= g.nz - 1)
    return true;  // the whole grid is scanned
  for (int s = 0; s < T; ++s) {
    const int rr = r - Q.dl[s];
    if (rr <= 0) return false;
    const double lim = (double)rr * g.h;
    if (!(bd[s] < lim * lim * (1.0 - 1e-9))) return false;
  }
  return true;
}

// the reference's result for a lexicographic minimum: 1-based element, or 0
// when no distance is below the reference's start value 1e14
__host__ __device__ inline int nearest_result(double bd, int be) {
  return (be != kNoElem && bd < 1e14) ? be + 1 : 0;
}

}  // namespace bcfd
