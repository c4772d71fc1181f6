// One full red-black SOR iteration (REDBLACKSOR_linear, PcorVcor.F90:313-349)
// per launch, streaming HBM once: each workgroup owns a TI x TJ column of
// interior cells over KS planes and marches it in k with the pc planes
// staged in LDS.
//
// Step s of the march (planes in a ring of four LDS slots), phases separated
// by workgroup barriers:
//   A: store plane s+1 (loaded into registers D steps earlier), issue the
//      loads of plane s+1+D, of the b/dt values step s+1 needs and of the
//      Acz row of plane s+1 (Acx, Acy of the tile sit in LDS for the march);
//   B: red update of plane s on the tile plus a one-cell halo (old black);
//   C: black update of plane s-1 on the tile (new red of planes s-2..s) and
//      write-back of plane s-1 of the tile, both colours, to the destination.
// The halo red values are recomputed by every tile that needs them from the
// same old black values, so they equal the owner's.  Source and destination
// arrays alternate between iterations (the old black values of a tile's halo
// must survive until every neighbour has read them).
//
// Slot layout: per plane and colour a (TJ+4) x (TI+4)/2 array holding slot
// row y = j - jL, half-column h = (i - iL) >> 1, where (iL, jL) is the tile
// origin minus two.  TI and TJ are even, so the colour of slot cell (x, y) in
// plane k is (x + y + k) & 1 (1 = red, as (i + j + k) odd).
//
// Index work per step is 32-bit: every thread keeps the slot positions of its
// cells (computed once) and offsets within one plane; the plane's base
// pointer is uniform.
//
// The phase functions are the per-thread bodies of the kernel; a host build
// of this header runs the same bodies thread by thread, phase by phase, to
// check the schedule (bounds and bit-identity with the in-place sweep).
//
// Real = double is the reference's SOR; Real = float the smoothing sweeps of
// the mixed-precision solve (same schedule, FP32 iterate and rhs).
#pragma once

#include "biocfd_rt.hpp"

namespace bcfd {

// colour-split half-row index of cell (h, j, k): h = (i - 1) >> 1
__host__ __device__ inline int64_t sidx(int NXh, int NY, int h, int j, int k) {
  return (int64_t)h + (int64_t)NXh * ((int64_t)(j - 1) + (int64_t)NY * (int64_t)(k - 1));
}

// right-hand side of a cell that the SOR does not update (not interior, or
// not fluid): a NaN with a private payload, tested bitwise
constexpr unsigned long long kNotFluid = 0x7ff4dead0000beefull;
constexpr unsigned kNotFluid32 = 0x7fa0beefu;
__host__ __device__ inline bool sor_skip(double v) {
  return __builtin_bit_cast(unsigned long long, v) == kNotFluid;
}
__host__ __device__ inline bool sor_skip(float v) { return __builtin_bit_cast(unsigned, v) == kNotFluid32; }
template <typename Real>
__host__ __device__ inline Real not_fluid_marker();
template <>
__host__ __device__ inline double not_fluid_marker<double>() {
  return __builtin_bit_cast(double, kNotFluid);
}
template <>
__host__ __device__ inline float not_fluid_marker<float>() {
  return __builtin_bit_cast(float, kNotFluid32);
}
__host__ __device__ inline double not_fluid_marker() { return not_fluid_marker<double>(); }

// Real = double: the reference's SOR on pc with rhs b/dt (bit-identical);
// Real = float: the smoothing sweeps of the mixed-precision solve, on the
// correction e with the FP64 residual rounded to FP32 as rhs.
template <typename Real>
struct FusedArgsT {
  Dims d;
  int NXh;                 // (NX + 1) / 2, row length of the colour-split rhs arrays
  int KS;                  // planes per workgroup
  int zero_src;            // src is all zero (first sweep from a zero guess): not read
  const Real *src;         // iterate at the start of the iteration (full layout)
  Real *dst;               // iterate after the iteration (full layout)
  const Real *bred;        // rhs of red cells, colour-split (marker where skipped)
  const Real *bblack;      // rhs of black cells
  const double *Ax, *Ay, *Az;
  double omega;
};
using FusedArgs = FusedArgsT<double>;

// stencil coefficients of one axis at one index: west/south/bottom, centre,
// east/north/top (columns 1..3 of Acx/Acy/Acz, CoefficientMatrix.f90)
template <typename Real>
struct Coef3T {
  Real m, c, p;
};
using Coef3 = Coef3T<double>;

// the SOR update of one cell: operand order of PcorVcor.F90:318-327
template <typename Real>
__host__ __device__ inline Real sor_cell(Real omega, const Coef3T<Real> &ax, const Coef3T<Real> &ay,
                                         const Coef3T<Real> &az, Real rhs, Real old, Real pw, Real pe, Real ps,
                                         Real pn, Real pb, Real pt) {
  const Real gs = (rhs - ax.m * pw - ax.p * pe - ay.m * ps - ay.p * pn - az.m * pb - az.p * pt) /
                  (ax.c + ay.c + az.c);
  return (Real(1) - omega) * old + omega * gs;
}

// coefficients of axis index q (2-based interior index, n interior cells)
template <typename Real = double>
__host__ __device__ inline Coef3T<Real> coef_at(const double *A, int n, int q) {
  return Coef3T<Real>{Real(A[q - 2]), Real(A[(q - 2) + n]), Real(A[(q - 2) + 2 * n])};
}

template <typename Real, int TI, int TJ, int NT, int D>
struct SorTile {
  static_assert(TI % 2 == 0 && TJ % 2 == 0, "tile sides must be even");
  static_assert(TI <= 500 && TJ <= 250, "slot positions are packed in 8 bits");
  static_assert(D == 1 || D == 2, "prefetch depth is 1 or 2 planes");
  static constexpr int PW = TI + 4;                 // slot row width in cells
  static constexpr int SH = TJ + 4;                 // slot rows
  static constexpr int SW = PW / 2;                 // half-row per colour
  static constexpr int PLANE = SH * PW;             // cells of one slot plane
  static constexpr int NL = (PLANE + NT - 1) / NT;  // plane loads per thread
  static constexpr int RW = TI / 2 + 1;             // red cells per row of the tile + halo
  static constexpr int NRC = (TJ + 2) * RW;         // red cells per plane
  static constexpr int NR = (NRC + NT - 1) / NT;
  static constexpr int HB = TI / 2;                 // red/black cell pairs per tile row
  static constexpr int NBC = TJ * HB;               // pairs per plane
  static constexpr int NB = (NBC + NT - 1) / NT;
  static constexpr int SLOTS = 4 * 2 * SH * SW;
  static constexpr int CXW = TI + 2, CYW = TJ + 2;  // staged coefficients per axis column
  static constexpr int LDS_DOUBLES = SLOTS + 3 * CXW + 3 * CYW;

  // where a workgroup's tile sits
  struct Tile {
    int iL, jL, k0, k1;
  };

  using Args = FusedArgsT<Real>;
  using C3 = Coef3T<Real>;
  // the update norm of the reference's convergence test; the FP32 sweeps
  // of the mixed solve are judged by the FP64 residual instead
  static constexpr bool kNorm = sizeof(Real) == sizeof(double);

  // per-thread registers of the march
  struct Thread {
    Real pcn[D][NL];          // planes in flight
    Real rn[NR], rc[NR];      // rhs of the red cells: next step, this step
    Real bn[NB], bc[NB];      // rhs of the black cells: next step, this step
    int ld_off[NL];           // in-plane offset of the loaded slot cell, -1 outside the array
    int ld_lds[NL];           // 2 * (y * SW + (x >> 1)) + ((x + y) & 1), -1 past the slot
    int red_yr[NR];           // y << 8 | r: row y, r-th red cell of the row; -1 past the end
    int pair_yh[NB];          // y << 8 | h of a red/black pair; -1 past the end
    C3 czm, cz0, czp;         // Acz of planes s-1, s, s+1
  };

  __host__ __device__ static Real *slot(Real *lds, int k, int colour) {
    return lds + ((k & 3) * 2 + colour) * (SH * SW);
  }

  // tile coordinates of linear workgroup L of n: consecutive tiles go to one
  // XCD (workgroups are dispatched to the 8 XCDs round-robin), so the halo
  // rows two neighbouring tiles share are fetched into the same L2
  __host__ __device__ static int xcd_tile(int L, int n) {
    const int per = n / 8, rem = n % 8, xcd = L % 8, q = L / 8;
    return xcd < rem ? xcd * (per + 1) + q : rem * (per + 1) + (xcd - rem) * per + q;
  }

  __host__ __device__ static Tile tile_of(const Args &a, int t, int gx, int gy) {
    const int tx = t % gx, ty = (t / gx) % gy, tz = t / (gx * gy);
    Tile T;
    T.iL = TI * tx;
    T.jL = TJ * ty;
    T.k0 = 2 + a.KS * tz;
    T.k1 = T.k0 + a.KS - 1 < a.d.nz + 1 ? T.k0 + a.KS - 1 : a.d.nz + 1;
    return T;
  }

  __host__ __device__ static C3 cz_of(const Args &a, int k) {
    if (k < 2 || k > a.d.nz + 1) return C3{Real(0), Real(0), Real(0)};
    return coef_at<Real>(a.Az, a.d.nz, k);
  }

  __host__ __device__ static void init_thread(const Args &a, const Tile &T, int tid, Thread &th) {
    for (int q = 0; q < NL; ++q) {
      const int e = tid + q * NT;
      th.ld_off[q] = -1;
      th.ld_lds[q] = -1;
      if (e < PLANE) {
        const int y = e / PW, x = e - y * PW;
        const int i = T.iL + x, j = T.jL + y;
        th.ld_lds[q] = 2 * (y * SW + (x >> 1)) + ((x + y) & 1);
        if (i >= 1 && i <= a.d.NX && j >= 1 && j <= a.d.NY) th.ld_off[q] = (i - 1) + a.d.NX * (j - 1);
      }
    }
    for (int q = 0; q < NR; ++q) {
      const int e = tid + q * NT;
      th.red_yr[q] = -1;
      if (e < NRC) {
        const int y = 1 + e / RW;
        th.red_yr[q] = (y << 8) | (e - (y - 1) * RW);
      }
    }
    for (int q = 0; q < NB; ++q) {
      const int e = tid + q * NT;
      th.pair_yh[q] = -1;
      if (e < NBC) {
        const int y = 2 + e / HB;
        th.pair_yh[q] = (y << 8) | (1 + (e - (y - 2) * HB));
      }
    }
    th.czm = cz_of(a, T.k0 - 2);
    th.cz0 = cz_of(a, T.k0 - 1);
    th.czp = th.cz0;
  }

  // Acx over slot columns x = 1..TI+2 and Acy over slot rows y = 1..TJ+2
  // (zero outside the interior, where no cell is updated)
  __host__ __device__ static void load_coef(const Args &a, const Tile &T, Real *lds, int tid) {
    Real *cx = lds + SLOTS, *cy = cx + 3 * CXW;
    for (int e = tid; e < 3 * CXW + 3 * CYW; e += NT) {
      if (e < 3 * CXW) {
        const int m = e / CXW, i = T.iL + 1 + (e - m * CXW);
        cx[e] = (i >= 2 && i <= a.d.nx + 1) ? Real(a.Ax[(i - 2) + m * a.d.nx]) : Real(0);
      } else {
        const int f = e - 3 * CXW, m = f / CYW, j = T.jL + 1 + (f - m * CYW);
        cy[f] = (j >= 2 && j <= a.d.ny + 1) ? Real(a.Ay[(j - 2) + m * a.d.ny]) : Real(0);
      }
    }
  }
  __host__ __device__ static C3 cx_at(const Real *lds, int x) {
    const Real *cx = lds + SLOTS;
    return C3{cx[x - 1], cx[CXW + x - 1], cx[2 * CXW + x - 1]};
  }
  __host__ __device__ static C3 cy_at(const Real *lds, int y) {
    const Real *cy = lds + SLOTS + 3 * CXW;
    return C3{cy[y - 1], cy[CYW + y - 1], cy[2 * CYW + y - 1]};
  }

  // plane p of src over the slot footprint (zero outside the array)
  __host__ __device__ static void load_plane(const Args &a, const Thread &th, int p, Real *r) {
    const bool pin = p >= 1 && p <= a.d.NZ && !a.zero_src;
    const Real *P = a.src + (int64_t)(pin ? p - 1 : 0) * ((int64_t)a.d.NX * a.d.NY);
    for (int q = 0; q < NL; ++q) r[q] = (pin && th.ld_off[q] >= 0) ? P[th.ld_off[q]] : Real(0);
  }

  __host__ __device__ static void store_plane(Real *lds, const Thread &th, int p, const Real *r) {
    Real *S = slot(lds, p, 0);
    for (int q = 0; q < NL; ++q) {
      const int l = th.ld_lds[q];
      if (l >= 0) S[((l & 1) ^ (p & 1)) * (SH * SW) + (l >> 1)] = r[q];
    }
  }

  // red cell of packed position yr in plane s: slot row y, half-column h, x = 2h + off
  __host__ __device__ static void red_cell(int yr, int s, int &y, int &h, int &x) {
    y = yr >> 8;
    const int off = (1 - y - s) & 1;
    h = 1 - off + (yr & 255);
    x = 2 * h + off;
  }

  __host__ __device__ static void load_bdt_red(const Args &a, const Tile &T, const Thread &th, int s, Real *r) {
    const bool kin = s >= 2 && s <= a.d.nz + 1;
    const Real *P = a.bred + (int64_t)(kin ? s - 1 : 0) * ((int64_t)a.NXh * a.d.NY);
    for (int q = 0; q < NR; ++q) {
      Real v = not_fluid_marker<Real>();
      if (kin && th.red_yr[q] >= 0) {
        int y, h, x;
        red_cell(th.red_yr[q], s, y, h, x);
        const int i = T.iL + x, j = T.jL + y;
        if (i >= 2 && i <= a.d.nx + 1 && j >= 2 && j <= a.d.ny + 1) v = P[((i - 1) >> 1) + a.NXh * (j - 1)];
      }
      r[q] = v;
    }
  }

  // pair of packed position yh in plane k: the black cell is at x = 2h + off,
  // the red one at x = 2h + 1 - off
  __host__ __device__ static void pair_cell(int yh, int k, int &y, int &h, int &off) {
    y = yh >> 8;
    h = yh & 255;
    off = (0 - y - k) & 1;
  }

  __host__ __device__ static void load_bdt_black(const Args &a, const Tile &T, const Thread &th, int k, Real *r) {
    const Real *P = a.bblack + (int64_t)(k - 1) * ((int64_t)a.NXh * a.d.NY);
    for (int q = 0; q < NB; ++q) {
      Real v = not_fluid_marker<Real>();
      if (th.pair_yh[q] >= 0) {
        int y, h, off;
        pair_cell(th.pair_yh[q], k, y, h, off);
        const int i = T.iL + 2 * h + off, j = T.jL + y;
        if (i <= a.d.nx + 1 && j <= a.d.ny + 1) v = P[((i - 1) >> 1) + a.NXh * (j - 1)];
      }
      r[q] = v;
    }
  }

  // red update of plane s (tile + halo); returns the largest |change|
  __host__ __device__ static double red_phase(const Args &a, Real *lds, const Thread &th, int s) {
    Real *R = slot(lds, s, 1);
    const Real *B = slot(lds, s, 0), *Bm = slot(lds, s - 1, 0), *Bp = slot(lds, s + 1, 0);
    double delta = 0.0;
    for (int q = 0; q < NR; ++q) {
      if (th.red_yr[q] < 0 || sor_skip(th.rc[q])) continue;
      int y, h, x;
      red_cell(th.red_yr[q], s, y, h, x);
      const int c = y * SW + h;
      const Real old = R[c];
      const Real nw = sor_cell<Real>(Real(a.omega), cx_at(lds, x), cy_at(lds, y), th.cz0, th.rc[q], old,
                                     B[y * SW + ((x - 1) >> 1)], B[y * SW + ((x + 1) >> 1)], B[c - SW], B[c + SW],
                                     Bm[c], Bp[c]);
      R[c] = nw;
      if (kNorm) {
        const double dd = fabs(double(nw - old));
        delta = delta > dd ? delta : dd;
      }
    }
    return delta;
  }

  // black update of plane k (tile) and write-back of the tile's plane k
  __host__ __device__ static double black_phase(const Args &a, const Tile &T, Real *lds, const Thread &th, int k) {
    const Real *R = slot(lds, k, 1), *Rm = slot(lds, k - 1, 1), *Rp = slot(lds, k + 1, 1);
    const Real *B = slot(lds, k, 0);
    Real *P = a.dst + (int64_t)(k - 1) * ((int64_t)a.d.NX * a.d.NY);
    double delta = 0.0;
    for (int q = 0; q < NB; ++q) {
      if (th.pair_yh[q] < 0) continue;
      int y, h, off;
      pair_cell(th.pair_yh[q], k, y, h, off);
      const int j = T.jL + y;
      if (j > a.d.ny + 1) continue;
      const int x = 2 * h + off, i = T.iL + x, ir = T.iL + 2 * h + 1 - off;
      const int c = y * SW + h;
      const int row = a.d.NX * (j - 1) - 1;
      if (i <= a.d.nx + 1) {
        const Real old = B[c];
        Real nw = old;
        if (!sor_skip(th.bc[q])) {
          nw = sor_cell<Real>(Real(a.omega), cx_at(lds, x), cy_at(lds, y), th.czm, th.bc[q], old,
                              R[y * SW + ((x - 1) >> 1)], R[y * SW + ((x + 1) >> 1)], R[c - SW], R[c + SW], Rm[c],
                              Rp[c]);
          if (kNorm) {
            const double dd = fabs(double(nw - old));
            delta = delta > dd ? delta : dd;
          }
        }
        P[row + i] = nw;
      }
      if (ir <= a.d.nx + 1) P[row + ir] = R[c];
    }
    return delta;
  }

  // ---- the march: prologue, then per step s = k0-1 .. k1+1 the phases A | B | C ----
  __host__ __device__ static void prologue(const Args &a, const Tile &T, Real *lds, Thread &th) {
    Real *r = th.pcn[0];
    load_plane(a, th, T.k0 - 2, r);
    store_plane(lds, th, T.k0 - 2, r);
    load_plane(a, th, T.k0 - 1, r);
    store_plane(lds, th, T.k0 - 1, r);
    for (int d = 0; d < D; ++d) load_plane(a, th, T.k0 + d, th.pcn[d]);
    load_bdt_red(a, T, th, T.k0 - 1, th.rn);
    for (int q = 0; q < NB; ++q) th.bn[q] = Real(0);
  }

  // buf holds plane s+1 on entry and plane s+1+D on exit
  __host__ __device__ static void phase_a(const Args &a, const Tile &T, Real *lds, Thread &th, int s, Real *buf) {
    store_plane(lds, th, s + 1, buf);
    for (int q = 0; q < NR; ++q) th.rc[q] = th.rn[q];
    for (int q = 0; q < NB; ++q) th.bc[q] = th.bn[q];
    if (s + 1 + D <= T.k1 + 2) load_plane(a, th, s + 1 + D, buf);
    if (s + 1 <= T.k1 + 1) load_bdt_red(a, T, th, s + 1, th.rn);
    if (s >= T.k0 && s <= T.k1) load_bdt_black(a, T, th, s, th.bn);
    th.czp = cz_of(a, s + 1);
  }

  __host__ __device__ static double phase_b(const Args &a, Real *lds, const Thread &th, int s) {
    return red_phase(a, lds, th, s);
  }

  __host__ __device__ static double phase_c(const Args &a, const Tile &T, Real *lds, Thread &th, int s) {
    const double d = (s - 1 >= T.k0) ? black_phase(a, T, lds, th, s - 1) : 0.0;
    th.czm = th.cz0;
    th.cz0 = th.czp;
    return d;
  }
};

}  // namespace bcfd
