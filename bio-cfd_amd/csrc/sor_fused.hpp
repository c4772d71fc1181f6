// NIT (1 or 2) full red-black SOR iterations (REDBLACKSOR_linear,
// PcorVcor.F90:313-349) per launch, streaming HBM once: each workgroup owns a
// TI x TJ column of interior cells over KS planes and marches it in k with the
// pc planes staged in a ring of 2*NIT+2 LDS slots.
//
// Step s of the march, phases separated by workgroup barriers:
//   A: store plane s+1 and its Acz row (loaded into registers D steps
//      earlier), issue the loads of plane s+1+D with its Acz row and of the
//      rhs values step s+1 needs (Acx, Acy of the tile sit in LDS for the
//      march; the Acz rows in a ring of R rows beside the plane slots);
//   p = 0 .. 2*NIT-1: update colour red (p even) or black (p odd) of
//      iteration p/2+1 on plane s-p over the tile plus a halo of
//      2*NIT-1-p cells (x, y and z); the last phase (black of the last
//      iteration, no halo) writes plane s-p of the tile, both colours, to
//      the destination.
// With NIT = 1: red on plane s (tile + 1), black on plane s-1 (tile).
// Every phase reads the other colour's values of the stage it needs, in
// place in LDS: a red update of iteration t reads black of iteration t-1 on
// planes that black of iteration t has not reached yet, and so on.  Halo
// values are recomputed by every tile that needs them from the same inputs,
// so they equal the owner's.  Source and destination arrays alternate between
// launches (the old values of a tile's halo must survive until every
// neighbour has read them).
//
// Slot layout: per plane and colour a (TJ+2H) x (TI+2H)/2 array (H = 2*NIT)
// holding slot row y = j - jL, half-column h = (i - iL) >> 1, where (iL, jL)
// is the tile origin minus H.  TI, TJ and H are even, so the colour of slot
// cell (x, y) in plane k is (x + y + k) & 1 (1 = red, as (i + j + k) odd).
//
// Index work per step is 32-bit: every thread keeps the slot positions of its
// cells (computed once) and offsets within one plane; the plane's base
// pointer is uniform.
//
// No value loaded from global memory is used in the step that loads it: the
// loads of a step complete in order, so waiting on one (a uniform Acz row,
// say) would also wait on the plane prefetched before it.
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

template <typename Real, int TI, int TJ, int NT, int D, int NIT = 1>
struct SorTile {
  static_assert(TI % 2 == 0 && TJ % 2 == 0, "tile sides must be even");
  static_assert(TI <= 500 && TJ <= 250, "slot positions are packed in 8 bits");
  static_assert(D == 1 || D == 2, "prefetch depth is 1 or 2 planes");
  static_assert(NIT == 1 || NIT == 2, "one or two iterations per pass");
  static constexpr int H = 2 * NIT;                 // slot halo (cells) around the tile
  static constexpr int NPH = 2 * NIT;               // update phases per step: red, black (, red, black)
  static constexpr int R = 2 * NIT + 2;             // planes in the LDS ring
  static constexpr int PW = TI + 2 * H;             // slot row width in cells
  static constexpr int SH = TJ + 2 * H;             // slot rows
  static constexpr int SW = PW / 2;                 // half-row per colour
  static constexpr int PLANE = SH * PW;             // cells of one slot plane
  static constexpr int NL = (PLANE + NT - 1) / NT;  // plane loads per thread
  static constexpr int SLOTS = R * 2 * SH * SW;
  static constexpr int CXW = PW - 2, CYW = SH - 2;  // staged coefficients: slot x (y) = 1 .. PW-2 (SH-2)
  static constexpr int CZ = SLOTS + 3 * CXW + 3 * CYW;  // Acz ring: 3 per slot plane
  static constexpr int SINK = CZ + 3 * R;                // written by slots past a phase's region
  static constexpr int LDS_DOUBLES = SINK + 1;

  // phase p (0 .. NPH-1) updates colour red (p even) or black (p odd) of
  // iteration p/2 + 1 on plane s - p over the tile plus a halo of hp(p)
  // cells in x, y and z
  __host__ __device__ static constexpr int hp(int p) { return 2 * NIT - 1 - p; }
  __host__ __device__ static constexpr int RWp(int p) { return TI / 2 + hp(p); }   // cells per row
  __host__ __device__ static constexpr int NCp(int p) { return (TJ + 2 * hp(p)) * RWp(p); }
  __host__ __device__ static constexpr int NQp(int p) { return (NCp(p) + NT - 1) / NT; }
  static constexpr int NQ = (NCp(0) + NT - 1) / NT;  // the largest phase

  // where a workgroup's tile sits: slot cell (x, y) of plane k is cell
  // (iL + x, jL + y, k); the tile's own cells have x, y in [H, H + TI (TJ))
  // and k in [k0, k1]
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
    double czn[D];            // their Acz rows: thread m < 3 carries column m
    Real rc[NPH][NQ];         // rhs of each phase's cells for this step
    Real rn[NIT == 1 ? NPH : 1][NIT == 1 ? NQ : 1];  // NIT = 1: the next step's, loaded in phase A
    int ld_off[NL];           // in-plane offset of the loaded slot cell, -1 outside the array
    int ld_lds[NL];           // 2 * (y * SW + (x >> 1)) + ((x + y) & 1), -1 past the slot
    int pos[NPH][NQ];         // y << 8 | r: row y (from the phase's first row), r-th cell; -1 past the end
    int czm;                  // Acz column this thread carries (0 .. 2), -1 for none
  };

  __host__ __device__ static int slot_index(int k) { return (k + 2 * R) % R; }
  __host__ __device__ static Real *slot(Real *lds, int k, int colour) {
    return lds + (slot_index(k) * 2 + colour) * (SH * SW);
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
    T.iL = 2 + TI * tx - H;
    T.jL = 2 + TJ * ty - H;
    T.k0 = 2 + a.KS * tz;
    T.k1 = T.k0 + a.KS - 1 < a.d.nz + 1 ? T.k0 + a.KS - 1 : a.d.nz + 1;
    return T;
  }

  // the march runs steps s = first_step .. last_step
  __host__ __device__ static int first_step(const Tile &T) { return T.k0 - (2 * NIT - 1); }
  __host__ __device__ static int last_step(const Tile &T) { return T.k1 + (2 * NIT - 1); }
  // phase p works on plane s - p at step s when that plane is in its range
  __host__ __device__ static bool active(const Tile &T, int p, int s) {
    return s - p >= T.k0 - hp(p) && s - p <= T.k1 + hp(p);
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
    for (int p = 0; p < NPH; ++p)
      for (int q = 0; q < NQ; ++q) {
        const int e = tid + q * NT;
        th.pos[p][q] = -1;
        if (q < NQp(p) && e < NCp(p)) {
          const int yy = e / RWp(p);
          th.pos[p][q] = (yy << 8) | (e - yy * RWp(p));
        }
      }
    th.czm = tid < 3 ? tid : -1;
  }

  // Acx over slot columns x = 1..PW-2 and Acy over slot rows y = 1..SH-2
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

  // plane p of src over the slot footprint (zero outside the array) and
  // column czm of its Acz row (zero off the interior, where no cell is
  // updated)
  __host__ __device__ static void load_plane(const Args &a, const Thread &th, int p, Real *r, double &cz) {
    const bool pin = p >= 1 && p <= a.d.NZ && !a.zero_src;
    const Real *P = a.src + (int64_t)(pin ? p - 1 : 0) * ((int64_t)a.d.NX * a.d.NY);
    for (int q = 0; q < NL; ++q) r[q] = (pin && th.ld_off[q] >= 0) ? P[th.ld_off[q]] : Real(0);
    cz = (th.czm >= 0 && p >= 2 && p <= a.d.nz + 1) ? a.Az[(p - 2) + th.czm * a.d.nz] : 0.0;
  }

  __host__ __device__ static void store_plane(Real *lds, const Thread &th, int p, const Real *r, double cz) {
    Real *S = slot(lds, p, 0);
    for (int q = 0; q < NL; ++q) {
      const int l = th.ld_lds[q];
      if (l >= 0) S[((l & 1) ^ (p & 1)) * (SH * SW) + (l >> 1)] = r[q];
    }
    if (th.czm >= 0) lds[CZ + 3 * slot_index(p) + th.czm] = Real(cz);
  }
  __host__ __device__ static C3 cz_at(const Real *lds, int k) {
    const Real *c = lds + CZ + 3 * slot_index(k);
    return C3{c[0], c[1], c[2]};
  }

  // cell of packed position v in phase p on plane k: slot row y, column x
  // (colour red for even p, black for odd p)
  __host__ __device__ static void cell_of(int p, int v, int k, int &y, int &x) {
    const int lo = H - hp(p);  // first slot row and column of the phase's region
    y = lo + (v >> 8);
    const int colour = (p & 1) ? 0 : 1;
    const int off = (colour - y - k) & 1;  // parity of the colour's columns in this row
    x = lo + ((off - lo) & 1) + 2 * (v & 255);
  }

  __host__ __device__ static void load_rhs(const Args &a, const Tile &T, const Thread &th, int p, int k,
                                           Real *r) {
    const bool kin = k >= 2 && k <= a.d.nz + 1;
    const Real *B = (p & 1) ? a.bblack : a.bred;
    const Real *P = B + (int64_t)(kin ? k - 1 : 0) * ((int64_t)a.NXh * a.d.NY);
    for (int q = 0; q < NQp(p); ++q) {
      Real v = not_fluid_marker<Real>();
      if (kin && th.pos[p][q] >= 0) {
        int y, x;
        cell_of(p, th.pos[p][q], k, y, x);
        const int i = T.iL + x, j = T.jL + y;
        if (i >= 2 && i <= a.d.nx + 1 && j >= 2 && j <= a.d.ny + 1) v = P[((i - 1) >> 1) + a.NXh * (j - 1)];
      }
      r[q] = v;
    }
  }

  // phase p at step s: update its colour on plane k = s - p (in LDS), and in
  // the last phase write plane k of the tile, both colours, to dst; returns
  // the largest |change| when norm is set.  The cells of all slots are
  // computed without branches (a slot past the region works on the region's
  // first cell and writes to the sink; a skipped cell keeps its value), so
  // their LDS reads are in flight together.
  __host__ __device__ static double update(const Args &a, const Tile &T, Real *lds, const Thread &th, int p, int s,
                                           bool norm) {
    const int k = s - p;
    const int own = (p & 1) ? 0 : 1, other = 1 - own;
    Real *O = slot(lds, k, own);
    const Real *N = slot(lds, k, other), *Nm = slot(lds, k - 1, other), *Np = slot(lds, k + 1, other);
    const bool last = p == NPH - 1;
    const C3 cz = cz_at(lds, k);
    constexpr int NQM = NQ;
    Real old[NQM], nw[NQM];
    int cs[NQM], xs[NQM], ys[NQM];
    for (int q = 0; q < NQp(p); ++q) {
      int y, x;
      cell_of(p, th.pos[p][q] < 0 ? 0 : th.pos[p][q], k, y, x);
      const int c = y * SW + (x >> 1);
      cs[q] = c;
      xs[q] = x;
      ys[q] = y;
      old[q] = O[c];
      const Real g = sor_cell<Real>(Real(a.omega), cx_at(lds, x), cy_at(lds, y), cz, th.rc[p][q], old[q],
                                    N[y * SW + ((x - 1) >> 1)], N[y * SW + ((x + 1) >> 1)], N[c - SW], N[c + SW],
                                    Nm[c], Np[c]);
      nw[q] = sor_skip(th.rc[p][q]) ? old[q] : g;
    }
    double delta = 0.0;
    if (kNorm && norm)
      for (int q = 0; q < NQp(p); ++q) {
        const double dd = th.pos[p][q] < 0 ? 0.0 : fabs(double(nw[q] - old[q]));
        delta = delta > dd ? delta : dd;
      }
    if (!last) {
      const int base = (int)(O - lds);
      for (int q = 0; q < NQp(p); ++q) lds[th.pos[p][q] < 0 ? SINK : base + cs[q]] = nw[q];
    } else {
      Real *P = a.dst + (int64_t)(k - 1) * ((int64_t)a.d.NX * a.d.NY);
      for (int q = 0; q < NQp(p); ++q) {
        if (th.pos[p][q] < 0) continue;
        // black cell x and its red partner x ^ 1 (same half-column)
        const int i = T.iL + xs[q], ir = T.iL + (xs[q] ^ 1), j = T.jL + ys[q];
        if (j <= a.d.ny + 1) {
          const int row = a.d.NX * (j - 1) - 1;
          if (i <= a.d.nx + 1) P[row + i] = nw[q];
          if (ir <= a.d.nx + 1) P[row + ir] = N[cs[q]];
        }
      }
    }
    return delta;
  }

  // ---- the march: prologue, then per step s the phases A | 0 | 1 | ... | NPH-1 ----
  __host__ __device__ static void prologue(const Args &a, const Tile &T, Real *lds, Thread &th) {
    const int s0 = first_step(T);
    Real *r = th.pcn[0];
    double &c = th.czn[0];
    load_plane(a, th, s0 - 1, r, c);
    store_plane(lds, th, s0 - 1, r, c);
    load_plane(a, th, s0, r, c);
    store_plane(lds, th, s0, r, c);
    for (int d = 0; d < D; ++d) load_plane(a, th, s0 + 1 + d, th.pcn[d], th.czn[d]);
    // phase p first works on plane k0 - hp(p) (at step k0 - hp(p) + p)
    if (NIT == 1) {
      // the first step's rhs of phase p go to rn (phase A moves them to rc)
      for (int p = 0; p < NPH; ++p) {
        if (active(T, p, s0))
          load_rhs(a, T, th, p, s0 - p, th.rn[p]);
        else
          for (int q = 0; q < NQ; ++q) th.rn[p][q] = Real(0);
      }
    } else {
      for (int p = 0; p < NPH; ++p) load_rhs(a, T, th, p, T.k0 - hp(p), th.rc[p]);
    }
  }

  // phase A of step s: buffer d holds plane s+1 on entry and plane s+1+D on
  // exit
  __host__ __device__ static void phase_a(const Args &a, const Tile &T, Real *lds, Thread &th, int s, int d) {
    Real *buf = th.pcn[d];
    store_plane(lds, th, s + 1, buf, th.czn[d]);
    if (NIT == 1)
      for (int p = 0; p < NPH; ++p)
        for (int q = 0; q < NQ; ++q) th.rc[p][q] = th.rn[NIT == 1 ? p : 0][NIT == 1 ? q : 0];
    if (s + 1 + D <= last_step(T) + 1) load_plane(a, th, s + 1 + D, buf, th.czn[d]);
    if (NIT == 1)
      for (int p = 0; p < NPH; ++p)
        if (s + 1 <= last_step(T) && active(T, p, s + 1)) load_rhs(a, T, th, p, s + 1 - p, th.rn[NIT == 1 ? p : 0]);
  }

  // phase p of step s (after phase A and phases 0 .. p-1).  With NIT = 1 the
  // next step's rhs were loaded in phase A (double-buffered); with NIT = 2
  // (twice the rhs registers) each phase reloads its own a step ahead of use
  __host__ __device__ static double phase_u(const Args &a, const Tile &T, Real *lds, Thread &th, int p, int s,
                                            bool norm) {
    if (!active(T, p, s)) return 0.0;
    const double d = update(a, T, lds, th, p, s, norm);
    if (NIT > 1 && active(T, p, s + 1)) load_rhs(a, T, th, p, s + 1 - p, th.rc[p]);
    return d;
  }
};

}  // namespace bcfd
