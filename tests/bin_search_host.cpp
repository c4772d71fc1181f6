// Host run of the binned nearest-element search of bio-cfd_amd/csrc/
// bin_search.hpp against the reference's scan (every element in order,
// strict <, start value 1e14; Search.F90:305-338, :762-858).
//
// The bins are built as the GPU builds them (bounding box, bin_grid_of,
// counts, exclusive scan, fill), each query runs the same lane bodies for the
// 64 lanes of a wavefront with the lane minima reduced after each shell, and
// the result must equal the reference scan's for every point.  Surfaces:
// sphere meshes of several sizes, a flat plate (one bin layer), duplicated
// centroids (exact ties: the lowest element must win), one element; query
// cells near the surface, inside the body and far outside the grid.  Built
// with AddressSanitizer, so any index outside the bin arrays aborts.
//
//   bin_search_host          exit 0 when every query matches
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bin_search.hpp"

using namespace bcfd;

static unsigned long long g_rng = 0x9e3779b97f4a7c15ull;
static double urand() {
  g_rng ^= g_rng << 13;
  g_rng ^= g_rng >> 7;
  g_rng ^= g_rng << 17;
  return (double)(g_rng >> 11) * (1.0 / 9007199254740992.0);
}

struct Bins {
  BinGrid g;
  std::vector<int> start, elem;
};

static Bins build(const std::vector<double> &x, const std::vector<double> &y, const std::vector<double> &z) {
  const int ne = (int)x.size();
  double mn[3] = {1e308, 1e308, 1e308}, mx[3] = {-1e308, -1e308, -1e308};
  for (int e = 0; e < ne; ++e) {
    const double v[3] = {x[e], y[e], z[e]};
    for (int a = 0; a < 3; ++a) {
      mn[a] = v[a] < mn[a] ? v[a] : mn[a];
      mx[a] = v[a] > mx[a] ? v[a] : mx[a];
    }
  }
  Bins B;
  B.g = bin_grid_of(mn, mx);
  const int nb = B.g.nx * B.g.ny * B.g.nz;
  if (nb > kMaxBins) {
    std::printf("bin grid %dx%dx%d exceeds kMaxBins\n", B.g.nx, B.g.ny, B.g.nz);
    std::exit(2);
  }
  std::vector<int> fill(nb, 0);
  for (int e = 0; e < ne; ++e) ++fill.at(bin_of(B.g, x[e], y[e], z[e]));
  B.start.assign(nb + 1, 0);
  for (int b = 0; b < nb; ++b) B.start[b + 1] = B.start[b] + fill[b];
  for (int b = 0; b < nb; ++b) fill[b] = B.start[b];
  B.elem.assign(ne, -1);
  // fill in reverse element order (the GPU's order within a bin is arbitrary)
  for (int e = ne - 1; e >= 0; --e) B.elem.at(fill[bin_of(B.g, x[e], y[e], z[e])]++) = e;
  return B;
}

// the reference's scan for one point
static int reference_scan(const std::vector<double> &x, const std::vector<double> &y, const std::vector<double> &z,
                          double px, double py, double pz) {
  double md = 1e14;
  int best = 0;
  for (size_t m = 0; m < x.size(); ++m) {
    const double ax = px - x[m], ay = py - y[m], az = pz - z[m];
    const double d = ax * ax + ay * ay + az * az;
    if (d < md) {
      md = d;
      best = (int)m + 1;
    }
  }
  return best;
}

// one wavefront's search, lane by lane
template <int T>
static void wave_search(const Bins &B, const std::vector<double> &x, const std::vector<double> &y,
                        const std::vector<double> &z, const double *px, const double *py, const double *pz, int *out,
                        int *shells) {
  // the bin arrays exactly as large as the GPU's (ASan checks the reads)
  std::vector<int> st(B.start), el(B.elem);
  const NearestQuery<T> Q = make_query<T>(B.g, px, py, pz);
  double bd[kSearchLanes][T];
  int be[kSearchLanes][T];
  for (int l = 0; l < kSearchLanes; ++l)
    for (int s = 0; s < T; ++s) {
      bd[l][s] = kNoDist;
      be[l][s] = kNoElem;
    }
  int r = 0;
  for (;; ++r) {
    for (int l = 0; l < kSearchLanes; ++l)
      shell_scan<T>(B.g, st.data(), el.data(), x.data(), y.data(), z.data(), Q, r, l, kSearchLanes, bd[l], be[l]);
    for (int s = 0; s < T; ++s) {
      double d = bd[0][s];
      int e = be[0][s];
      for (int l = 1; l < kSearchLanes; ++l)
        if (lex_less(bd[l][s], be[l][s], d, e)) {
          d = bd[l][s];
          e = be[l][s];
        }
      for (int l = 0; l < kSearchLanes; ++l) {
        bd[l][s] = d;
        be[l][s] = e;
      }
    }
    if (shell_done<T>(B.g, Q, r, bd[0])) break;
  }
  for (int s = 0; s < T; ++s) out[s] = nearest_result(bd[0][s], be[0][s]);
  *shells = r + 1;
}

// a cell of side h at (cx, cy, cz): the p point and six face points, as the
// normal-distance search uses them, and the low corner as retagging does
static int check_cell(const char *name, const Bins &B, const std::vector<double> &x, const std::vector<double> &y,
                      const std::vector<double> &z, double cx, double cy, double cz, double h, long &nq,
                      long &shell_sum) {
  const double px[7] = {cx, cx - h / 2, cx + h / 2, cx, cx, cx, cx};
  const double py[7] = {cy, cy, cy, cy - h / 2, cy + h / 2, cy, cy};
  const double pz[7] = {cz, cz, cz, cz, cz, cz - h / 2, cz + h / 2};
  int got[7], sh;
  wave_search<7>(B, x, y, z, px, py, pz, got, &sh);
  shell_sum += sh;
  int bad = 0;
  for (int s = 0; s < 7; ++s) {
    const int want = reference_scan(x, y, z, px[s], py[s], pz[s]);
    if (got[s] != want) {
      std::printf("MISMATCH %s point (%.17g, %.17g, %.17g): %d vs reference %d\n", name, px[s], py[s], pz[s], got[s],
                  want);
      ++bad;
    }
  }
  const double qx[2] = {cx, cx - h / 2}, qy[2] = {cy, cy - h / 2}, qz[2] = {cz, cz - h / 2};
  int got2[2];
  wave_search<2>(B, x, y, z, qx, qy, qz, got2, &sh);
  shell_sum += sh;
  for (int s = 0; s < 2; ++s) {
    const int want = reference_scan(x, y, z, qx[s], qy[s], qz[s]);
    if (got2[s] != want) {
      std::printf("MISMATCH %s (2-point) (%.17g, %.17g, %.17g): %d vs reference %d\n", name, qx[s], qy[s], qz[s],
                  got2[s], want);
      ++bad;
    }
  }
  nq += 9;
  return bad;
}

static void sphere(int ne, double r0, std::vector<double> &x, std::vector<double> &y, std::vector<double> &z) {
  x.resize(ne);
  y.resize(ne);
  z.resize(ne);
  const double ga = 3.14159265358979323846 * (3.0 - std::sqrt(5.0));
  for (int e = 0; e < ne; ++e) {
    const double t = 1.0 - 2.0 * (e + 0.5) / ne, rr = std::sqrt(1.0 - t * t);
    x[e] = 2.0 + r0 * rr * std::cos(ga * e);
    y[e] = 2.0 + r0 * rr * std::sin(ga * e);
    z[e] = 2.0 + r0 * t;
  }
}

static int run_surface(const char *name, const std::vector<double> &x, const std::vector<double> &y,
                       const std::vector<double> &z, double h, int ncells, double spread) {
  const Bins B = build(x, y, z);
  int bad = 0;
  long nq = 0, shells = 0;
  for (int c = 0; c < ncells; ++c) {
    // cells near random elements, some inside and some far away
    const int e = (int)(urand() * x.size()) % (int)x.size();
    const double cx = x[e] + spread * (urand() - 0.5), cy = y[e] + spread * (urand() - 0.5),
                 cz = z[e] + spread * (urand() - 0.5);
    bad += check_cell(name, B, x, y, z, cx, cy, cz, h, nq, shells);
  }
  // far outside the grid, and exactly at centroids (zero distances, ties)
  bad += check_cell(name, B, x, y, z, -50.0, 3.0, 2.0, h, nq, shells);
  bad += check_cell(name, B, x, y, z, 2.0, 2.0, 40.0, h, nq, shells);
  bad += check_cell(name, B, x, y, z, x[0], y[0], z[0], h, nq, shells);
  std::printf("%s %s: %zu elements, bins %dx%dx%d, %ld points, %.2f shells per query\n", bad ? "FAIL" : "ok", name,
              x.size(), B.g.nx, B.g.ny, B.g.nz, nq, (double)shells / (nq / 4.5));
  return bad;
}

int main() {
  int bad = 0;
  std::vector<double> x, y, z;
  // sphere d = 1 with 8192 elements, cells of 1/32 near the surface
  sphere(8192, 0.5, x, y, z);
  bad += run_surface("sphere8192", x, y, z, 1.0 / 32, 300, 4.0 / 32);
  bad += run_surface("sphere8192 wide", x, y, z, 1.0 / 32, 60, 1.5);
  sphere(1024, 0.1, x, y, z);
  bad += run_surface("sphere1024", x, y, z, 1.0 / 64, 200, 4.0 / 64);
  // flat plate: every centroid at z = 2 (one bin layer)
  {
    x.clear(); y.clear(); z.clear();
    for (int a = 0; a < 64; ++a)
      for (int b = 0; b < 64; ++b) {
        x.push_back(1.5 + a / 63.0);
        y.push_back(1.5 + b / 63.0);
        z.push_back(2.0);
      }
    bad += run_surface("plate", x, y, z, 1.0 / 32, 200, 6.0 / 32);
  }
  // duplicated centroids: exact ties between different elements
  {
    sphere(512, 0.3, x, y, z);
    const size_t n = x.size();
    for (size_t e = 0; e < n; ++e) {
      x.push_back(x[e]);
      y.push_back(y[e]);
      z.push_back(z[e]);
    }
    bad += run_surface("duplicates", x, y, z, 1.0 / 32, 200, 3.0 / 32);
  }
  // one element
  x.assign(1, 1.0);
  y.assign(1, 1.0);
  z.assign(1, 1.0);
  bad += run_surface("single", x, y, z, 0.1, 20, 1.0);
  std::printf(bad ? "FAILED %d\n" : "all queries equal the reference scan\n", bad);
  return bad ? 1 : 0;
}
