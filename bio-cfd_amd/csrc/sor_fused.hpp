// One full red-black SOR iteration (REDBLACKSOR_linear, PcorVcor.F90:313-349)
// per launch, streaming HBM once: each workgroup owns a TI x TJ column of
// interior cells over KS planes and marches it in k with the pc planes
// staged in LDS.
//
// Step s of the march (planes in a ring of four LDS slots):
//   store plane s+1 (prefetched into registers during step s-1),
//   issue the loads of plane s+2, of the b/dt values and of the Acz row that
//   step s+1 needs (Acx and Acy of the tile sit in LDS for the whole march),
//   red update of plane s on the tile plus a one-cell halo (reads old black),
//   black update of plane s-1 on the tile (reads the new red of planes
//   s-2..s) and write plane s-1 of the tile to the destination array.
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
// The functions below are the per-thread bodies of the kernel's phases; a
// host build of this header runs the same bodies thread by thread to check
// the schedule (bounds and bit-identity with the in-place sweep).
#pragma once

#include "biocfd_rt.hpp"

namespace bcfd {

// colour-split half-row index of cell (h, j, k): h = (i - 1) >> 1
__host__ __device__ inline int64_t sidx(int NXh, int NY, int h, int j, int k) {
  return (int64_t)h + (int64_t)NXh * ((int64_t)(j - 1) + (int64_t)NY * (int64_t)(k - 1));
}

// b/dt of a cell that the SOR does not update (not interior, or not fluid):
// a NaN with a private payload, tested bitwise
constexpr unsigned long long kNotFluid = 0x7ff4dead0000beefull;
__host__ __device__ inline bool sor_skip(double v) {
  return __builtin_bit_cast(unsigned long long, v) == kNotFluid;
}
__host__ __device__ inline double not_fluid_marker() { return __builtin_bit_cast(double, kNotFluid); }

struct FusedArgs {
  Dims d;
  int NXh;                 // (NX + 1) / 2, row length of the colour-split b/dt arrays
  int KS;                  // planes per workgroup
  const double *src;       // pc at the start of the iteration (full layout)
  double *dst;             // pc after the iteration (full layout)
  const double *bred;      // b/dt of red cells, colour-split (marker where skipped)
  const double *bblack;    // b/dt of black cells
  const double *Ax, *Ay, *Az;
  double omega;
};

// stencil coefficients of one axis at one index: west/south/bottom, centre,
// east/north/top (columns 1..3 of Acx/Acy/Acz, CoefficientMatrix.f90)
struct Coef3 {
  double m, c, p;
};

// the SOR update of one cell: operand order of PcorVcor.F90:318-327
__host__ __device__ inline double sor_cell(double omega, const Coef3 &ax, const Coef3 &ay, const Coef3 &az,
                                           double rhs, double old, double pw, double pe, double ps, double pn,
                                           double pb, double pt) {
  const double gs = (rhs - ax.m * pw - ax.p * pe - ay.m * ps - ay.p * pn - az.m * pb - az.p * pt) /
                    (ax.c + ay.c + az.c);
  return (1.0 - omega) * old + omega * gs;
}

// coefficients of axis index q (2-based interior index, n interior cells)
__host__ __device__ inline Coef3 coef_at(const double *A, int n, int q) {
  return Coef3{A[q - 2], A[(q - 2) + n], A[(q - 2) + 2 * n]};
}

template <int TI, int TJ, int NT>
struct SorTile {
  static_assert(TI % 2 == 0 && TJ % 2 == 0, "tile sides must be even");
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

  __host__ __device__ static double *slot(double *lds, int k, int colour) {
    return lds + ((k & 3) * 2 + colour) * (SH * SW);
  }

  // Acx over slot columns x = 1..TI+2 and Acy over slot rows y = 1..TJ+2
  // (zero outside the interior, where no cell is updated)
  __host__ __device__ static void load_coef(const FusedArgs &a, double *lds, int tid, int iL, int jL) {
    double *cx = lds + SLOTS, *cy = cx + 3 * CXW;
    for (int e = tid; e < 3 * CXW + 3 * CYW; e += NT) {
      if (e < 3 * CXW) {
        const int m = e / CXW, i = iL + 1 + (e - m * CXW);
        cx[e] = (i >= 2 && i <= a.d.nx + 1) ? a.Ax[(i - 2) + m * a.d.nx] : 0.0;
      } else {
        const int f = e - 3 * CXW, m = f / CYW, j = jL + 1 + (f - m * CYW);
        cy[f] = (j >= 2 && j <= a.d.ny + 1) ? a.Ay[(j - 2) + m * a.d.ny] : 0.0;
      }
    }
  }
  __host__ __device__ static Coef3 cx_at(const double *lds, int x) {
    const double *cx = lds + SLOTS;
    return Coef3{cx[x - 1], cx[CXW + x - 1], cx[2 * CXW + x - 1]};
  }
  __host__ __device__ static Coef3 cy_at(const double *lds, int y) {
    const double *cy = lds + SLOTS + 3 * CXW;
    return Coef3{cy[y - 1], cy[CYW + y - 1], cy[2 * CYW + y - 1]};
  }
  // Acz of plane k (zero outside the interior)
  __host__ __device__ static Coef3 cz_of(const FusedArgs &a, int k) {
    if (k < 2 || k > a.d.nz + 1) return Coef3{0.0, 0.0, 0.0};
    return coef_at(a.Az, a.d.nz, k);
  }

  // plane p of src over the slot footprint (zero outside the array)
  __host__ __device__ static void load_plane(const FusedArgs &a, int tid, int p, int iL, int jL, double *r) {
    const bool pin = p >= 1 && p <= a.d.NZ;
    for (int q = 0; q < NL; ++q) {
      const int e = tid + q * NT;
      double v = 0.0;
      if (pin && e < PLANE) {
        const int y = e / PW, x = e - y * PW;
        const int i = iL + x, j = jL + y;
        if (i >= 1 && i <= a.d.NX && j >= 1 && j <= a.d.NY) v = a.src[ix3(a.d, i, j, p)];
      }
      r[q] = v;
    }
  }

  __host__ __device__ static void store_plane(double *lds, int tid, int p, const double *r) {
    for (int q = 0; q < NL; ++q) {
      const int e = tid + q * NT;
      if (e < PLANE) {
        const int y = e / PW, x = e - y * PW;
        slot(lds, p, (x + y + p) & 1)[y * SW + (x >> 1)] = r[q];
      }
    }
  }

  // red cell e of plane s in the tile + halo: slot row y, half-column h, x = 2h + off
  __host__ __device__ static void red_cell(int e, int s, int &y, int &h, int &x) {
    y = 1 + e / RW;
    const int r = e - (y - 1) * RW;
    const int off = (1 - y - s) & 1;
    h = 1 - off + r;
    x = 2 * h + off;
  }

  __host__ __device__ static void load_bdt_red(const FusedArgs &a, int tid, int s, int iL, int jL, double *r) {
    for (int q = 0; q < NR; ++q) {
      const int e = tid + q * NT;
      double v = not_fluid_marker();
      if (e < NRC) {
        int y, h, x;
        red_cell(e, s, y, h, x);
        const int i = iL + x, j = jL + y;
        if (i >= 2 && i <= a.d.nx + 1 && j >= 2 && j <= a.d.ny + 1 && s >= 2 && s <= a.d.nz + 1)
          v = a.bred[sidx(a.NXh, a.d.NY, (i - 1) >> 1, j, s)];
      }
      r[q] = v;
    }
  }

  // pair e of plane k in the tile: slot row y, half-column h; the black cell
  // is at x = 2h + off, the red one at x = 2h + 1 - off
  __host__ __device__ static void pair_cell(int e, int k, int &y, int &h, int &off) {
    y = 2 + e / HB;
    h = 1 + (e - (y - 2) * HB);
    off = (0 - y - k) & 1;
  }

  __host__ __device__ static void load_bdt_black(const FusedArgs &a, int tid, int k, int iL, int jL, double *r) {
    for (int q = 0; q < NB; ++q) {
      const int e = tid + q * NT;
      double v = not_fluid_marker();
      if (e < NBC) {
        int y, h, off;
        pair_cell(e, k, y, h, off);
        const int i = iL + 2 * h + off, j = jL + y;
        if (i <= a.d.nx + 1 && j <= a.d.ny + 1) v = a.bblack[sidx(a.NXh, a.d.NY, (i - 1) >> 1, j, k)];
      }
      r[q] = v;
    }
  }

  // red update of plane s (tile + halo); returns the largest |change|
  __host__ __device__ static double red_phase(const FusedArgs &a, double *lds, int tid, int s, const Coef3 &cz,
                                              const double *rb) {
    double *R = slot(lds, s, 1);
    const double *B = slot(lds, s, 0), *Bm = slot(lds, s - 1, 0), *Bp = slot(lds, s + 1, 0);
    double delta = 0.0;
    for (int q = 0; q < NR; ++q) {
      const int e = tid + q * NT;
      if (e >= NRC || sor_skip(rb[q])) continue;
      int y, h, x;
      red_cell(e, s, y, h, x);
      const int c = y * SW + h;
      const double old = R[c];
      const double nw = sor_cell(a.omega, cx_at(lds, x), cy_at(lds, y), cz, rb[q], old, B[y * SW + ((x - 1) >> 1)],
                                 B[y * SW + ((x + 1) >> 1)], B[c - SW], B[c + SW], Bm[c], Bp[c]);
      R[c] = nw;
      const double dd = fabs(nw - old);
      delta = delta > dd ? delta : dd;
    }
    return delta;
  }

  // black update of plane k (tile) and write-back of the tile's plane k
  __host__ __device__ static double black_phase(const FusedArgs &a, double *lds, int tid, int k, int iL, int jL,
                                                const Coef3 &cz, const double *bb) {
    const double *R = slot(lds, k, 1), *Rm = slot(lds, k - 1, 1), *Rp = slot(lds, k + 1, 1);
    const double *B = slot(lds, k, 0);
    double delta = 0.0;
    for (int q = 0; q < NB; ++q) {
      const int e = tid + q * NT;
      if (e >= NBC) continue;
      int y, h, off;
      pair_cell(e, k, y, h, off);
      const int j = jL + y;
      if (j > a.d.ny + 1) continue;
      const int x = 2 * h + off, i = iL + x, ir = iL + 2 * h + 1 - off;
      const int c = y * SW + h;
      if (i <= a.d.nx + 1) {
        const double old = B[c];
        double nw = old;
        if (!sor_skip(bb[q])) {
          nw = sor_cell(a.omega, cx_at(lds, x), cy_at(lds, y), cz, bb[q], old, R[y * SW + ((x - 1) >> 1)],
                        R[y * SW + ((x + 1) >> 1)], R[c - SW], R[c + SW], Rm[c], Rp[c]);
          const double dd = fabs(nw - old);
          delta = delta > dd ? delta : dd;
        }
        a.dst[ix3(a.d, i, j, k)] = nw;
      }
      if (ir <= a.d.nx + 1) a.dst[ix3(a.d, ir, j, k)] = R[c];
    }
    return delta;
  }

  // tile coordinates of linear workgroup L of n: consecutive tiles go to one
  // XCD (workgroups are dispatched to the 8 XCDs round-robin), so the halo
  // rows two neighbouring tiles share are fetched into the same L2
  __host__ __device__ static int xcd_tile(int L, int n) {
    const int per = n / 8, rem = n % 8, xcd = L % 8, q = L / 8;
    return xcd < rem ? xcd * (per + 1) + q : rem * (per + 1) + (xcd - rem) * per + q;
  }
};

}  // namespace bcfd
