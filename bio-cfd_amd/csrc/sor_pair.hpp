// Pair-owned variant of the fused red-black SOR tile (NIT = 1 or 2 full
// iterations of REDBLACKSOR_linear, PcorVcor.F90:313-349, per launch): the
// same march, LDS ring, phases and halos as SorTile (sor_fused.hpp), with a
// thread-to-cell mapping that keeps per-cell work to the stencil itself.
//
// Every thread owns fixed pairs of cells (slot columns 2h and 2h+1 of one
// slot row y, 1 <= y <= SH-2) for the whole launch.  In a phase of colour c
// on plane k exactly one cell of each pair has that colour, the one at
// column 2h + par with par = (c ^ y ^ k) & 1, and both colours of a pair
// share the colour-split slot index y*SW + h.  So per phase a pair costs one
// parity bit: its slot address, its neighbours (W, E at index - 1 + par and
// + par of the other colour, S, N at -SW, +SW, B, T on the neighbour planes)
// and its coefficients follow from it.  With CREG the coefficients (Acx of
// the cell each step class updates, Acy of the row and the partial sums
// Acx(c) + Acy(c)) sit in registers from the start; without, Acx (even
// columns, then odd ones) and Acy of the tile are staged in LDS (fewer
// registers).  Whether the cell is updated (inside the phase's region and
// the interior in i and j) is a precomputed bit per (phase, step class); a
// pair whose cell is not updated writes its own value back, so the phase
// has no branches.  The
// pairs are numbered from the tile outwards, ring by ring of the shrinking
// phase regions, so the pairs a phase works on are a prefix of the
// numbering: a thread's pair slots past it drop out at compile time.
//
// One barrier per plane step.  Each thread loads (and stores into the LDS
// ring) the cells of its own pairs, plus a share of the two edge rows of the
// slot plane, so within a step every LDS word a thread reads is one it wrote
// itself or one written in an earlier step: the cell above a phase's cell
// (plane s+1 for phase 0, the cell phase p-1 just updated for phase p) is
// the same column of the same pair, read back from the thread's own store
// at the step's start (phase 0) or kept in a register; the cell below is
// the pair's own, updated by the thread's next phase; the four neighbours
// in the plane were written a step earlier.  So the phases of a step need
// no barrier between them, only the step boundary does (it orders a step's
// writes before the next step's reads and the ring slot's reuse).
// The rhs of step s+1's first two phases is loaded at the start of step s,
// plane s+2 right after phase 0, into registers indexed by the step
// number.  With two iterations per launch each cell's rhs is read once:
// phase p + 2 of step s + 2 updates the cells phase p updated at step s
// (same plane k = s - p, same colour, and, the pairs being numbered by
// phase region, the same threads and slots), so it reuses the registers
// loaded for phase p, which live four steps (buffers by step mod 4).
//
// Within a step the parity bit of every pair is the same for all phases (the
// colour of phase p and its plane k = s - p change together), and it flips
// from step to step.  So everything that depends on it is precomputed per
// step class (b = (s - first step) & 1, known at compile time in the
// unrolled step loop): the coefficients of the cell the class updates, the
// update bit of each phase and the parity itself.  The steps in the middle
// of the march, where every phase is active, every plane inside the block
// and the next step inside the march, run a copy of the step with none of
// these guards (groups of RB steps that fit entirely; the first and last
// groups keep them).  Plane loads go through a raw buffer resource over the
// plane: cells outside the array get an offset past its end, which the
// hardware range check returns as zero.
//
// Real = double is the reference's SOR (bit-identical: the operand order of
// sor_cell, Acx(c) + Acy(c) + Acz(c) summed left to right); Real = float the
// smoothing sweeps of the mixed-precision solve, which accumulate the
// stencil with fused multiply-adds and divide by the diagonal through the
// hardware reciprocal (v_rcp_f32, about 1 ulp): rounding-level changes of
// the FP32 smoother, whose corrections the FP64 residual of the defect
// correction checks.
#pragma once

#include "sor_fused.hpp"

namespace bcfd {

// 1/den in FP32: the hardware reciprocal on the GPU, 1/den on the host (the
// host run checks the schedule, not the reciprocal's last bit)
__host__ __device__ inline float sor_rcp(float den) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(den);
#else
  return 1.0f / den;
#endif
}

// the SOR update of one cell; den = (Acx + Acy) + Acz of the cell.  FP64:
// the operand order of PcorVcor.F90:318-327 (sor_cell), bit-identical.
// FP32: the stencil as a chain of fused multiply-adds (IEEE fma, the same
// bits on host and GPU), the reciprocal, and old + omega (gs - old).
__host__ __device__ inline double sor_update(double omega, double rhs, double old, double axm, double pw, double axp,
                                             double pe, double aym, double ps, double ayp, double pn, double azm,
                                             double pb, double azp, double pt, double den) {
  const double gs = (rhs - axm * pw - axp * pe - aym * ps - ayp * pn - azm * pb - azp * pt) / den;
  return (1.0 - omega) * old + omega * gs;
}
__host__ __device__ inline float sor_update(float omega, float rhs, float old, float axm, float pw, float axp,
                                            float pe, float aym, float ps, float ayp, float pn, float azm, float pb,
                                            float azp, float pt, float den) {
  float num = fmaf(-axm, pw, rhs);
  num = fmaf(-axp, pe, num);
  num = fmaf(-aym, ps, num);
  num = fmaf(-ayp, pn, num);
  num = fmaf(-azm, pb, num);
  num = fmaf(-azp, pt, num);
  const float gs = num * sor_rcp(den);
  return fmaf(omega, gs - old, old);
}

template <typename Real, int TI, int TJ, int NT, int NIT, bool CREG = true>
struct SorPairTile {
  static_assert(TI % 2 == 0 && TJ % 2 == 0, "tile sides must be even");
  static_assert(NIT == 1 || NIT == 2, "one or two iterations per pass");
  static constexpr int H = 2 * NIT;                 // slot halo (cells) around the tile
  static constexpr int NPH = 2 * NIT;               // update phases per step
  static constexpr int R = 2 * NIT + 2;             // planes in the LDS ring
  static constexpr int PW = TI + 2 * H, SH = TJ + 2 * H, SW = PW / 2;
  static constexpr int PLANE = SH * PW;             // cells of one slot plane
  static constexpr int NEDGE = 2 * PW;              // cells of slot rows 0 and SH-1 (no pairs)
  static constexpr int NPAIR = (SH - 2) * SW;       // pairs of slot rows 1 .. SH-2
  static constexpr int NQ = (NPAIR + NT - 1) / NT;  // pairs per thread
  static constexpr int NE = (NEDGE + NT - 1) / NT;  // edge cells per thread
  static constexpr int NL = 2 * NQ + NE;            // plane loads per thread: both cells of each pair, edges
  static constexpr int SLOTS = R * 2 * SH * SW;
  static constexpr int CZ = SLOTS;                  // Acz ring: 3 per slot plane
  static constexpr int SINK = CZ + 3 * R;           // written by load entries past the slot plane
  // staged Acx of every slot column, the even columns then the odd ones (a
  // wavefront's pairs, consecutive in h, read consecutive words), and Acy
  // of slot rows 1 .. SH-2
  static constexpr int CXW = PW, CYW = SH - 2;
  static constexpr int CX = SINK + 1, CY = CX + 3 * CXW;
  static constexpr int LDS_DOUBLES = CREG ? SINK + 1 : CY + 3 * CYW;
  // mask bits of a pair: b * NPH + p = in steps of class b (b = (s - first
  // step) & 1) phase p updates the pair's cell of that class; kW0, kW1 = the
  // last phase writes cell 2h, 2h+1 to the destination; kPar + b = the column
  // (2h + par) of the cell class b updates; bits 16-23: h, 24-31: y
  static constexpr int kW0 = 2 * NPH, kW1 = 2 * NPH + 1, kPar = 2 * NPH + 2;
  static_assert(kPar + 1 < 16 && SW < 256 && SH < 256, "pair mask layout");
  static constexpr unsigned kOut = 0x80000000u;     // byte offset of a cell outside the array
  static constexpr int NQC = CREG ? NQ : 1;         // pairs with coefficient registers
  static constexpr int RB = NIT == 2 ? 4 : 2;       // rhs register buffers (steps)

  // the pairs with a cell in phase p's region form the rectangle of slot
  // rows [ry0, ry1] and pairs [rh0, rh1]; numbered innermost (phase NPH-1,
  // the tile itself) first, so phase p works on pairs 0 .. cnt(p)-1, in
  // slots q < used(p) of its threads (slot q holds pair tid + q*NT)
  __host__ __device__ static constexpr int ry0(int p) { return p + 1; }
  __host__ __device__ static constexpr int ry1(int p) { return TJ + 4 * NIT - 2 - p; }
  __host__ __device__ static constexpr int rh0(int p) { return (p + 1) / 2; }
  __host__ __device__ static constexpr int rh1(int p) { return (TI + 4 * NIT - 2 - p) / 2; }
  __host__ __device__ static constexpr int cnt(int p) { return (ry1(p) - ry0(p) + 1) * (rh1(p) - rh0(p) + 1); }
  __host__ __device__ static constexpr int used(int p) { return (cnt(p) + NT - 1) / NT; }
  __host__ __device__ static constexpr bool full(int p, int q) { return (q + 1) * NT <= cnt(p); }
  static_assert(cnt(0) == NPAIR, "phase 0 covers every pair");

  // pair number e (0 <= e < NPAIR) -> pair h of slot row y
  __host__ __device__ static void pair_of(int e, int &h, int &y) {
    int p = NPH - 1;
    while (e >= cnt(p)) --p;  // the ring of level p: cnt(p+1) <= e < cnt(p)
    const int oy0 = ry0(p), oy1 = ry1(p), ox0 = rh0(p), ox1 = rh1(p), W = ox1 - ox0 + 1;
    if (p == NPH - 1) {
      y = oy0 + e / W;
      h = ox0 + e % W;
      return;
    }
    const int r = e - cnt(p + 1);
    const int iy0 = ry0(p + 1), iy1 = ry1(p + 1), ix0 = rh0(p + 1), ix1 = rh1(p + 1);
    const int nTop = (iy0 - oy0) * W, mL = ix0 - ox0, m = mL + (ox1 - ix1), nMid = (iy1 - iy0 + 1) * m;
    if (r < nTop) {
      y = oy0 + r / W;
      h = ox0 + r % W;
    } else if (r < nTop + nMid) {
      const int r2 = r - nTop, c = r2 % m;
      y = iy0 + r2 / m;
      h = c < mL ? ox0 + c : ix1 + 1 + (c - mL);
    } else {
      const int r3 = r - nTop - nMid;
      y = iy1 + 1 + r3 / W;
      h = ox0 + r3 % W;
    }
  }

  struct Tile {
    int iL, jL, k0, k1;
  };
  using Args = FusedArgsT<Real>;
  static constexpr bool kNorm = sizeof(Real) == sizeof(double);

  struct Thread {
    Real pcn[NL];            // the plane in flight (entry 2q + c: column 2h + c of pair slot q)
    Real tv[NQ];             // the cell above the next phase's cell of each pair
    double czn;              // its Acz row: thread m < 3 carries column m
    unsigned ld_off[NL];     // in-plane byte offset of the loaded slot cell, kOut outside the array
    int ld_lds[NL];          // LDS index of the slot cell in the ring slot of an even plane, -1 past the slot
    int czm;                 // Acz column this thread carries (0 .. 2), -1 for none
    int tid;
    int c[NQ];               // slot index y*SW + h of each pair
    unsigned mask[NQ];
    int roff[NQ];            // rhs in-plane offset of column 2h (column 2h+1: + 1)
    int goff[NQ];            // iterate in-plane offset of column 2h
    Real cxm[NQC][2], cxp[NQC][2], sxy[NQC][2], cym[NQC], cyp[NQC];  // [b]: the cell of step class b
    Real rc[RB][2][NQ];      // rhs of the cells of phases 0 and 1, by step mod RB
  };

  __host__ __device__ static int slot_index(int k) { return (k + 2 * R) % R; }
  __host__ __device__ static int slot_off(int k, int colour) { return (slot_index(k) * 2 + colour) * (SH * SW); }

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

  __host__ __device__ static constexpr int hp(int p) { return 2 * NIT - 1 - p; }
  __host__ __device__ static int first_step(const Tile &T) { return T.k0 - (2 * NIT - 1); }
  __host__ __device__ static int last_step(const Tile &T) { return T.k1 + (2 * NIT - 1); }
  __host__ __device__ static bool active(const Tile &T, int p, int s) {
    return s - p >= T.k0 - hp(p) && s - p <= T.k1 + hp(p);
  }

  // plane load entry: slot cell (x, y), its LDS index in the ring slot of an
  // even plane (colour of (x, y) there: (x + y) & 1) and its in-plane byte
  // offset, kOut outside the array
  __host__ __device__ static void load_entry(const Args &a, const Tile &T, int x, int y, int &lds_i, unsigned &off) {
    const int i = T.iL + x, j = T.jL + y;
    lds_i = ((x + y) & 1) * (SH * SW) + y * SW + (x >> 1);
    off = (i >= 1 && i <= a.d.NX && j >= 1 && j <= a.d.NY)
              ? (unsigned)((i - 1) + a.d.NX * (j - 1)) * (unsigned)sizeof(Real) : kOut;
  }

  __host__ __device__ static void init_thread(const Args &a, const Tile &T, int tid, Thread &th) {
    for (int q = 0; q < NL; ++q) {
      th.ld_off[q] = kOut;
      th.ld_lds[q] = -1;
    }
    for (int q = 0; q < NQ; ++q) {
      const int e = tid + q * NT;
      if (e >= NPAIR) continue;
      int h, y;
      pair_of(e, h, y);
      for (int c = 0; c < 2; ++c) load_entry(a, T, 2 * h + c, y, th.ld_lds[2 * q + c], th.ld_off[2 * q + c]);
    }
    for (int r = 0; r < NE; ++r) {
      const int e = tid + r * NT;
      if (e >= NEDGE) continue;
      load_entry(a, T, e % PW, e < PW ? 0 : SH - 1, th.ld_lds[2 * NQ + r], th.ld_off[2 * NQ + r]);
    }
    th.czm = tid < 3 ? tid : -1;
    th.tid = tid;
    const int nxh = a.NXh * a.d.NY;
    // the colour of phase p and its plane k = s - p give (colour ^ k) & 1 =
    // 1 ^ (s & 1): a pair updates column 2h + ((y ^ 1 ^ s) & 1) in every
    // phase of step s, and s & 1 = (first step & 1) ^ b in class b
    const int s0par = first_step(T) & 1;
    for (int q = 0; q < NQ; ++q) {
      const int e = tid + q * NT;
      const bool real = e < NPAIR;
      int y = 1, h = 0;
      if (real) pair_of(e, h, y);
      const int j = T.jL + y, i0 = T.iL + 2 * h;
      const bool jin = j >= 2 && j <= a.d.ny + 1;
      unsigned m = (unsigned)h << 16 | (unsigned)y << 24;
      for (int b = 0; b < 2; ++b) {
        const int par = (y ^ 1 ^ s0par ^ b) & 1;
        const int x = 2 * h + par, i = i0 + par;
        const bool iin = i >= 2 && i <= a.d.nx + 1;
        m |= (unsigned)par << (kPar + b);
        for (int p = 0; p < NPH; ++p) {
          const int lo = p + 1, hi = lo + 2 * hp(p);  // region: [lo, lo + T + 2 hp)
          if (real && iin && jin && x >= lo && x < hi + TI && y >= lo && y < hi + TJ) m |= 1u << (b * NPH + p);
        }
        if constexpr (CREG) {
          const Coef3T<Real> cx = iin ? coef_at<Real>(a.Ax, a.d.nx, i) : Coef3T<Real>{Real(0), Real(0), Real(0)};
          const Coef3T<Real> cy = jin ? coef_at<Real>(a.Ay, a.d.ny, j) : Coef3T<Real>{Real(0), Real(0), Real(0)};
          th.cxm[q][b] = cx.m;
          th.cxp[q][b] = cx.p;
          th.sxy[q][b] = cx.c + cy.c;
          th.cym[q] = cy.m;
          th.cyp[q] = cy.p;
        }
      }
      // the tile's own pairs: the last phase writes both cells
      const bool own = real && y >= H && y < H + TJ && h >= H / 2 && h < H / 2 + TI / 2 && j <= a.d.ny + 1;
      if (own && i0 <= a.d.nx + 1) m |= 1u << kW0;
      if (own && i0 + 1 <= a.d.nx + 1) m |= 1u << kW1;
      th.mask[q] = m;
      th.c[q] = y * SW + h;
      // rhs of column 2h + par: half-column (i0 + par - 1) >> 1 = i0/2 - 1 + par
      int ro = a.NXh * (j - 1) + i0 / 2 - 1;
      ro = ro < 0 ? 0 : (ro > nxh - 2 ? nxh - 2 : ro);
      th.roff[q] = ro;
      th.goff[q] = (i0 - 1) + a.d.NX * (j - 1);
    }
  }

  __host__ __device__ static bool plane_in(const Args &a, int p) { return p >= 1 && p <= a.d.NZ && !a.zero_src; }

  // the steps that need no guard: every phase active (s >= k0 + 2 NIT - 1),
  // every phase's plane inside the block (2 NIT + 1 <= s <= nz + 1), the
  // planes s+1 and s+2 inside the array and step s+1 inside the march
  // (s <= min(last step - 1, nz)); none when the source is not read
  __host__ __device__ static int steady_lo(const Args &, const Tile &T) {
    return T.k0 + 2 * NIT - 1 > 2 * NIT + 1 ? T.k0 + 2 * NIT - 1 : 2 * NIT + 1;
  }
  __host__ __device__ static int steady_hi(const Args &a, const Tile &T) {
    if (a.zero_src) return -(1 << 30);
    return last_step(T) - 1 < a.d.nz ? last_step(T) - 1 : a.d.nz;
  }
  // the group of RB steps starting at s runs without guards
  __host__ __device__ static bool group_steady(const Args &a, const Tile &T, int s) {
    return s >= steady_lo(a, T) && s + RB - 1 <= steady_hi(a, T);
  }

  // plane p of src over the slot footprint (cells outside the array read as
  // zero) and column czm of its Acz row
  template <bool S>
  __host__ __device__ static void load_plane(const Args &a, Thread &th, int p) {
    if (S || plane_in(a, p)) {
      const Real *P = a.src + (int64_t)(p - 1) * ((int64_t)a.d.NX * a.d.NY);
      const unsigned nbytes = (unsigned)a.d.NX * (unsigned)a.d.NY * (unsigned)sizeof(Real);
#if defined(__HIP_DEVICE_COMPILE__)
      const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(const_cast<Real *>(P), (short)0, (int)nbytes,
                                                                         0x00020000);
      for (int q = 0; q < NL; ++q) {
        if constexpr (sizeof(Real) == 8)
          th.pcn[q] = __builtin_bit_cast(Real, __builtin_amdgcn_raw_buffer_load_b64(r, (int)th.ld_off[q], 0, 0));
        else
          th.pcn[q] = __builtin_bit_cast(Real, __builtin_amdgcn_raw_buffer_load_b32(r, (int)th.ld_off[q], 0, 0));
      }
#else
      for (int q = 0; q < NL; ++q) th.pcn[q] = th.ld_off[q] >= nbytes ? Real(0) : P[th.ld_off[q] / sizeof(Real)];
#endif
    }
    th.czn = (th.czm >= 0 && p >= 2 && p <= a.d.nz + 1) ? a.Az[(p - 2) + th.czm * a.d.nz] : 0.0;
  }

  // the value plane p's load entry q stores
  template <bool S>
  __host__ __device__ static Real plane_value(const Args &a, const Thread &th, int p, int q) {
    return (S || plane_in(a, p)) ? th.pcn[q] : Real(0);
  }

  template <bool S>
  __host__ __device__ static void store_plane(const Args &a, Real *lds, const Thread &th, int p) {
    const int base = slot_off(p, 0), flip = (p & 1) * (SH * SW);
    for (int q = 0; q < NL; ++q) {
      const int l = th.ld_lds[q];
      // an odd plane swaps the colour halves
      const int w = l < 0 ? SINK : base + (l >= SH * SW ? l - flip : l + flip);
      lds[w] = plane_value<S>(a, th, p, q);
    }
    if (th.czm >= 0) lds[CZ + 3 * slot_index(p) + th.czm] = Real(th.czn);
  }

  // the rhs of phases 0 and 1 active at step s (class CL), into buffer rb
  // (phases 2 and 3 of step s + 2 reuse them)
  template <int CL, bool S>
  __host__ __device__ static void load_rhs(const Args &a, const Tile &T, Thread &th, int s, Real (*rb)[NQ]) {
    const int64_t plane = (int64_t)a.NXh * a.d.NY;
    for (int p = 0; p < 2; ++p) {
      if (!S && !active(T, p, s)) continue;
      const int k = s - p, colour = (p & 1) ? 0 : 1;
      const bool kin = S || (k >= 2 && k <= a.d.nz + 1);
      const Real *B = (colour ? a.bred : a.bblack) + (kin ? k - 1 : 1) * plane;
      for (int q = 0; q < NQ; ++q) {
        if (q >= used(p)) break;
        const int par = (int)((th.mask[q] >> (kPar + CL)) & 1u);
        rb[p][q] = B[th.roff[q] + par];
      }
    }
  }

  // without CREG: Acx over every slot column (even columns, then odd) and
  // Acy over slot rows 1..SH-2 into LDS (zero outside the interior, where no
  // cell is updated)
  __host__ __device__ static void load_coef(const Args &a, const Tile &T, Real *lds, int tid) {
    if constexpr (!CREG) {
      for (int e = tid; e < 3 * CXW + 3 * CYW; e += NT) {
        if (e < 3 * CXW) {
          const int m = e / CXW, r = e - m * CXW, i = T.iL + (r < SW ? 2 * r : 2 * (r - SW) + 1);
          lds[CX + e] = (i >= 2 && i <= a.d.nx + 1) ? Real(a.Ax[(i - 2) + m * a.d.nx]) : Real(0);
        } else {
          const int f = e - 3 * CXW, m = f / CYW, j = T.jL + 1 + (f - m * CYW);
          lds[CY + f] = (j >= 2 && j <= a.d.ny + 1) ? Real(a.Ay[(j - 2) + m * a.d.ny]) : Real(0);
        }
      }
    }
  }

  // planes s0-1 and s0 into LDS, plane s0+1 and the rhs of step s0 in flight
  // (tid: the thread's index, for the coefficient staging)
  __host__ __device__ static void prologue(const Args &a, const Tile &T, Real *lds, Thread &th, int tid) {
    load_coef(a, T, lds, tid);
    const int s0 = first_step(T);
    load_plane<false>(a, th, s0 - 1);
    store_plane<false>(a, lds, th, s0 - 1);
    load_plane<false>(a, th, s0);
    store_plane<false>(a, lds, th, s0);
    load_plane<false>(a, th, s0 + 1);
    load_rhs<0, false>(a, T, th, s0, th.rc[0]);
  }

  // start of step s (rhs buffer B holds step s's): store plane s+1 (the
  // thread's own cells), issue the loads of step s+1's rhs into buffer B+1.
  // S: a step of a guard-free group
  template <int B, bool S = false>
  __host__ __device__ static void step_begin(const Args &a, const Tile &T, Real *lds, Thread &th, int s) {
    store_plane<S>(a, lds, th, s + 1);
    if (S || s + 1 <= last_step(T)) load_rhs<(B + 1) & 1, S>(a, T, th, s + 1, th.rc[(B + 1) % RB]);
  }

  // phase p of step s: update its colour on plane k = s - p (in LDS; the last
  // phase writes plane k of the tile, both colours, to dst); returns the
  // largest |change| when norm is set
  template <int B, bool S = false>
  __host__ __device__ static double phase_u(const Args &a, const Tile &T, Real *lds, Thread &th, int p, int s,
                                            bool norm) {
    const double delta = phase_cells<B, S>(a, T, lds, th, p, s, norm);
    // the loads of plane s+2 are issued after phase 0, not with the step's
    // start, so that their registers are not held through phase 0
    if (p == 0 && (S || s + 2 <= last_step(T) + 1)) load_plane<S>(a, th, s + 2);
    return delta;
  }

  template <int B, bool S>
  __host__ __device__ static double phase_cells(const Args &a, const Tile &T, Real *lds, Thread &th, int p, int s,
                                                bool norm) {
    if (!S && !active(T, p, s)) return 0.0;
    constexpr int CL = B & 1;  // the step's class (RB is even)
    const int k = s - p, colour = (p & 1) ? 0 : 1;
    const bool kin = S || (k >= 2 && k <= a.d.nz + 1);
    const bool last = p == NPH - 1;
    const int oO = slot_off(k, colour), oN = slot_off(k, 1 - colour);
    const int oNm = slot_off(k - 1, 1 - colour), oNp = slot_off(k + 1, 1 - colour);
    const Real *cz = lds + CZ + 3 * slot_index(k);
    const Real czm = cz[0], czc = cz[1], czp = cz[2];
    const Real omega = Real(a.omega);
    Real *P = a.dst + (int64_t)(k - 1) * ((int64_t)a.d.NX * a.d.NY);
    double delta = 0.0;
    for (int q = 0; q < NQ; ++q) {
      if (q >= used(p)) break;
      if (!full(p, q) && th.tid + q * NT >= cnt(p)) continue;
      const unsigned m = th.mask[q];
      const int par = (int)((m >> (kPar + CL)) & 1u);
      const int c = th.c[q];
      const Real *N = lds + oN + c;
      const Real old = lds[oO + c];
      const Real pw = N[par - 1], pe = N[par], ps = N[-SW], pn = N[SW];
      // above: plane s+1 as this thread stored it at the step's start (phase
      // 0), else the cell this thread's previous phase updated; below: the
      // pair's own cell
      const Real pb = lds[oNm + c];
      const Real pt = p == 0 ? lds[oNp + c] : th.tv[q];
      // phases 2, 3: the rhs phase p - 2 loaded two steps earlier
      const Real rhs = p < 2 ? th.rc[B][p < 2 ? p : 0][q] : th.rc[(B + RB - 2) % RB][p >= 2 ? p - 2 : 0][q];
      Real axm, axp, aym, ayp, sxy;
      if constexpr (CREG) {
        axm = th.cxm[q][CL];
        axp = th.cxp[q][CL];
        sxy = th.sxy[q][CL];
        aym = th.cym[q];
        ayp = th.cyp[q];
      } else {
        const Real *cx = lds + CX + par * SW + (int)((m >> 16) & 255u);  // column 2h + par
        const Real *cy = lds + CY + (int)(m >> 24) - 1;
        axm = cx[0];
        axp = cx[2 * CXW];
        aym = cy[0];
        ayp = cy[2 * CYW];
        sxy = cx[CXW] + cy[CYW];
      }
      // the centre sum (Acx + Acy) + Acz, as sor_cell's
      const Real nw = sor_update(omega, rhs, old, axm, pw, axp, pe, aym, ps, ayp, pn, czm, pb, czp, pt, sxy + czc);
      const bool upd = kin && ((m >> (CL * NPH + p)) & 1u) && !sor_skip(rhs);
      if (kNorm && norm) {
        const double dd = upd ? fabs(double(nw - old)) : 0.0;
        delta = delta > dd ? delta : dd;
      }
      const Real v = upd ? nw : old;
      if (!last) {
        // a cell that is not updated gets its own value back: only this
        // thread writes it, and it holds that value since an earlier step
        lds[oO + c] = v;
        th.tv[q] = v;
      } else {
        const Real partner = par ? pw : pe;  // the other colour in the same half-column
        if (m & (1u << kW0)) P[th.goff[q]] = par ? partner : v;
        if (m & (1u << kW1)) P[th.goff[q] + 1] = par ? v : partner;
      }
    }
    return delta;
  }
};

}  // namespace bcfd
