// Host run of the fused SOR tile schedule of bio-cfd_amd/csrc/sor_fused.hpp.
//
// The kernel's phases (plane store + prefetch, red update, black update) are
// executed thread by thread, workgroup by workgroup, with the same per-thread
// bodies the gfx950 kernel runs and the same barrier placement, on grids whose
// sides are not multiples of the tiles.  Built with AddressSanitizer on the
// host, so any index outside the arrays or the LDS image aborts.  The result
// of every iteration must equal, bit for bit, the textbook in-place red-black
// sweep (all red cells, then all black cells) on the same arrays.
//
//   sor_tile_host            exit 0 when every case matches
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <vector>

#include "sor_fused.hpp"
#include "sor_pair.hpp"

using namespace bcfd;

template <typename Real>
struct Problem {
  Dims d;
  int NXh;
  std::vector<Real> pc, bred, bblack;
  std::vector<double> Ax, Ay, Az;
  std::vector<unsigned char> fluid;
  double omega;
};

static unsigned long long g_rng = 88172645463325252ull;
static long long steady_steps = 0;  // pair-tile steps emulated in guard-free groups
static double urand() {
  g_rng ^= g_rng << 13;
  g_rng ^= g_rng >> 7;
  g_rng ^= g_rng << 17;
  return (double)(g_rng >> 11) * (1.0 / 9007199254740992.0);
}

template <typename Real>
static Problem<Real> make_problem(int nx, int ny, int nz, double solid_frac) {
  Problem<Real> p;
  p.d = Dims{nx, ny, nz, nx + 2, ny + 2, nz + 2, nx + 3, ny + 3};
  p.NXh = (p.d.NX + 1) / 2;
  const int64_t n = (int64_t)p.d.NX * p.d.NY * p.d.NZ;
  const int64_t nh = (int64_t)p.NXh * p.d.NY * p.d.NZ;
  p.pc.resize(n);
  p.fluid.assign(n, 0);
  p.bred.assign(nh, not_fluid_marker<Real>());
  p.bblack.assign(nh, not_fluid_marker<Real>());
  for (auto &v : p.pc) v = Real(urand() - 0.5);
  auto axis = [](std::vector<double> &A, int m) {
    A.resize(3 * (size_t)m);
    for (int q = 0; q < m; ++q) {
      const double h = 0.8 + 0.4 * urand();
      A[q] = 1.0 / (h * h);
      A[q + 2 * m] = 1.0 / (h * h) * (0.9 + 0.2 * urand());
      A[q + m] = -(A[q] + A[q + 2 * m]);
    }
  };
  axis(p.Ax, nx);
  axis(p.Ay, ny);
  axis(p.Az, nz);
  for (int k = 2; k <= nz + 1; ++k)
    for (int j = 2; j <= ny + 1; ++j)
      for (int i = 2; i <= nx + 1; ++i) {
        if (urand() < solid_frac) continue;
        p.fluid[ix3(p.d, i, j, k)] = 1;
        const Real v = Real((urand() - 0.5) * 4.0);
        auto &dst = ((i + j + k) & 1) ? p.bred : p.bblack;
        dst[sidx(p.NXh, p.d.NY, (i - 1) >> 1, j, k)] = v;
      }
  p.omega = 1.7;
  return p;
}

template <typename Real>
static FusedArgsT<Real> args_of(const Problem<Real> &p, int KS) {
  FusedArgsT<Real> a;
  a.d = p.d;
  a.NXh = p.NXh;
  a.KS = KS;
  a.zero_src = 0;
  a.src = nullptr;
  a.dst = nullptr;
  a.bred = p.bred.data();
  a.bblack = p.bblack.data();
  a.Ax = p.Ax.data();
  a.Ay = p.Ay.data();
  a.Az = p.Az.data();
  a.omega = p.omega;
  return a;
}

// textbook in-place red-black iteration; returns max |change|.  With RCP the
// update is sor_update's (the pair-owned tiles': FP64 identical to
// sor_cell, FP32 with fused multiply-adds and the reciprocal)
template <typename Real, bool RCP = false>
static double inplace_iteration(const Problem<Real> &p, std::vector<Real> &pc) {
  const FusedArgsT<Real> a = args_of(p, 1);
  const Dims &d = p.d;
  const int64_t sy = d.NX, sz = (int64_t)d.NX * d.NY;
  double delta = 0.0;
  for (int colour = 1; colour >= 0; --colour)
    for (int k = 2; k <= d.nz + 1; ++k)
      for (int j = 2; j <= d.ny + 1; ++j)
        for (int i = 2; i <= d.nx + 1; ++i) {
          if (((i + j + k) & 1) != colour) continue;
          const auto &bt = colour ? p.bred : p.bblack;
          const Real rhs = bt[sidx(p.NXh, d.NY, (i - 1) >> 1, j, k)];
          if (sor_skip(rhs)) continue;
          const int64_t c = ix3(d, i, j, k);
          const Real old = pc[c];
          const Coef3T<Real> ax = coef_at<Real>(a.Ax, d.nx, i), ay = coef_at<Real>(a.Ay, d.ny, j),
                             az = coef_at<Real>(a.Az, d.nz, k);
          Real nw;
          if (RCP) {
            nw = sor_update(Real(a.omega), rhs, old, ax.m, pc[c - 1], ax.p, pc[c + 1], ay.m, pc[c - sy], ay.p,
                            pc[c + sy], az.m, pc[c - sz], az.p, pc[c + sz], ax.c + ay.c + az.c);
          } else {
            nw = sor_cell<Real>(Real(a.omega), ax, ay, az, rhs, old, pc[c - 1], pc[c + 1], pc[c - sy], pc[c + sy],
                                pc[c - sz], pc[c + sz]);
          }
          pc[c] = nw;
          const double dd = std::fabs(double(nw - old));
          delta = delta > dd ? delta : dd;
        }
  return delta;
}

// one fused launch (NIT iterations) src -> dst, emulating every workgroup and
// thread: the kernel's step loop with each barrier-separated phase run for all
// threads; returns the update norm of the launch's last iteration
template <typename Real, int TI, int TJ, int NT, int D, int NIT>
static double fused_launch(const Problem<Real> &p, int KS, const std::vector<Real> &src, std::vector<Real> &dst,
                           bool zero_src) {
  using T = SorTile<Real, TI, TJ, NT, D, NIT>;
  FusedArgsT<Real> a = args_of(p, KS);
  a.src = src.data();
  a.dst = dst.data();
  a.zero_src = zero_src ? 1 : 0;
  const int gx = (p.d.nx + TI - 1) / TI, gy = (p.d.ny + TJ - 1) / TJ, gz = (p.d.nz + KS - 1) / KS;
  const int n = gx * gy * gz;
  std::vector<Real> lds(T::LDS_DOUBLES);
  std::vector<typename T::Thread> th(NT);
  std::vector<int> seen(n, 0);
  double delta = 0.0;
  for (int L = 0; L < n; ++L) {
    const int t = T::xcd_tile(L, n);
    if (t < 0 || t >= n || seen[t]++) {
      std::printf("xcd_tile is not a bijection at %d of %d\n", L, n);
      std::exit(2);
    }
    const typename T::Tile tile = T::tile_of(a, t, gx, gy);
    for (auto &v : lds) v = Real(std::nan(""));
    for (int tid = 0; tid < NT; ++tid) {
      T::init_thread(a, tile, tid, th[tid]);
      T::load_coef(a, tile, lds.data(), tid);
      T::prologue(a, tile, lds.data(), th[tid]);
    }
    const int s1 = T::last_step(tile);
    for (int s = T::first_step(tile); s <= s1; s += D)
      for (int d = 0; d < D; ++d) {
        if (d > 0 && s + d > s1) break;
        for (int tid = 0; tid < NT; ++tid) T::phase_a(a, tile, lds.data(), th[tid], s + d, d);
        for (int ph = 0; ph < T::NPH; ++ph)
          for (int tid = 0; tid < NT; ++tid) {  // barrier before each phase
            const double dd = T::phase_u(a, tile, lds.data(), th[tid], ph, s + d, ph >= T::NPH - 2);
            delta = delta > dd ? delta : dd;
          }
      }
  }
  return delta;
}

// the same launch with the pair-owned tile (sor_pair.hpp): the rhs buffer of
// a step alternates with the step's parity from the first step
template <typename Real, int TI, int TJ, int NT, int NIT, bool CREG>
static double pair_launch(const Problem<Real> &p, int KS, const std::vector<Real> &src, std::vector<Real> &dst,
                          bool zero_src, int order, int check = 1) {
  using T = SorPairTile<Real, TI, TJ, NT, NIT, CREG>;
  {
    // the pair numbering is a bijection onto slot rows 1..SH-2 x pairs
    // 0..SW-1, and the pairs of phase p (a prefix) have a cell in its region
    std::vector<int> hit((size_t)T::SH * T::SW, 0);
    for (int e = 0; e < T::NPAIR; ++e) {
      int h, y;
      T::pair_of(e, h, y);
      int lvl = T::NPH - 1;
      while (e >= T::cnt(lvl)) --lvl;
      const bool in = y >= 1 && y <= T::SH - 2 && h >= 0 && h < T::SW && y >= T::ry0(lvl) && y <= T::ry1(lvl) &&
                      h >= T::rh0(lvl) && h <= T::rh1(lvl);
      if (!in || hit[(size_t)y * T::SW + h]++) {
        std::printf("pair numbering broken at %d -> (%d, %d)\n", e, h, y);
        std::exit(2);
      }
    }
  }
  FusedArgsT<Real> a = args_of(p, KS);
  a.src = src.data();
  a.dst = dst.data();
  a.zero_src = zero_src ? 1 : 0;
  const int gx = (p.d.nx + TI - 1) / TI, gy = (p.d.ny + TJ - 1) / TJ, gz = (p.d.nz + KS - 1) / KS;
  const int n = gx * gy * gz;
  std::vector<Real> lds(T::LDS_DOUBLES);
  std::vector<typename T::Thread> th(NT);
  double delta = 0.0;
  for (int L = 0; L < n; ++L) {
    const typename T::Tile tile = T::tile_of(a, T::xcd_tile(L, n), gx, gy);
    for (auto &v : lds) v = Real(std::nan(""));
    for (int tid = 0; tid < NT; ++tid) {
      T::init_thread(a, tile, tid, th[tid]);
      T::prologue(a, tile, lds.data(), th[tid], tid);
    }
    const int s0 = T::first_step(tile), s1 = T::last_step(tile);
    for (int s = s0; s <= s1; ++s) {
      const int b = (s - s0) % T::RB;  // the kernel's step unrolling
      // the kernel's guard-free groups of RB steps
      const bool st = T::group_steady(a, tile, s - b);
      steady_steps += st;
      auto begin = [&](int tid) {
        switch (b + (st ? 4 : 0)) {
          case 0: T::template step_begin<0, false>(a, tile, lds.data(), th[tid], s); break;
          case 1: T::template step_begin<1, false>(a, tile, lds.data(), th[tid], s); break;
          case 2: T::template step_begin<2 % T::RB, false>(a, tile, lds.data(), th[tid], s); break;
          case 3: T::template step_begin<3 % T::RB, false>(a, tile, lds.data(), th[tid], s); break;
          case 4: T::template step_begin<0, true>(a, tile, lds.data(), th[tid], s); break;
          case 5: T::template step_begin<1, true>(a, tile, lds.data(), th[tid], s); break;
          case 6: T::template step_begin<2 % T::RB, true>(a, tile, lds.data(), th[tid], s); break;
          default: T::template step_begin<3 % T::RB, true>(a, tile, lds.data(), th[tid], s); break;
        }
      };
      auto phase = [&](int tid, int ph) {
        // the norm of the launch's last iteration (check 1) or of its first
        // (check 2), as pair_step selects the phases
        const bool nm = (check == 1 && ph >= T::NPH - 2) || (check == 2 && ph < 2);
        double dd;
        switch (b + (st ? 4 : 0)) {
          case 0: dd = T::template phase_u<0, false>(a, tile, lds.data(), th[tid], ph, s, nm); break;
          case 1: dd = T::template phase_u<1, false>(a, tile, lds.data(), th[tid], ph, s, nm); break;
          case 2: dd = T::template phase_u<2 % T::RB, false>(a, tile, lds.data(), th[tid], ph, s, nm); break;
          case 3: dd = T::template phase_u<3 % T::RB, false>(a, tile, lds.data(), th[tid], ph, s, nm); break;
          case 4: dd = T::template phase_u<0, true>(a, tile, lds.data(), th[tid], ph, s, nm); break;
          case 5: dd = T::template phase_u<1, true>(a, tile, lds.data(), th[tid], ph, s, nm); break;
          case 6: dd = T::template phase_u<2 % T::RB, true>(a, tile, lds.data(), th[tid], ph, s, nm); break;
          default: dd = T::template phase_u<3 % T::RB, true>(a, tile, lds.data(), th[tid], ph, s, nm); break;
        }
        delta = delta > dd ? delta : dd;
      };
      // the step's only barrier is at its start: any interleaving of the
      // threads within a step must give the same result; order 0 runs the
      // phases one after the other over all threads, 1 and 2 each thread's
      // whole step, threads ascending / descending
      if (order == 0) {
        for (int tid = 0; tid < NT; ++tid) begin(tid);
        for (int ph = 0; ph < T::NPH; ++ph)
          for (int tid = 0; tid < NT; ++tid) phase(tid, ph);
      } else {
        for (int n = 0; n < NT; ++n) {
          const int tid = order == 1 ? n : NT - 1 - n;
          begin(tid);
          for (int ph = 0; ph < T::NPH; ++ph) phase(tid, ph);
        }
      }
    }
  }
  return delta;
}

// `launches` fused launches (NIT iterations each) against the in-place
// sweep; with zero_first the first starts from a zero iterate without reading
// it (the first smoothing pass of a mixed-precision correction)
// Pair: 0 = SorTile, 1 = SorPairTile with coefficient registers, 2 = with
// coefficients in LDS
template <typename Real, int TI, int TJ, int NT, int D, int NIT, int Pair = 0>
static int run_case(int nx, int ny, int nz, int KS, int launches, bool zero_first, int order = 0) {
  Problem<Real> p = make_problem<Real>(nx, ny, nz, 0.1);
  std::vector<Real> ref = p.pc, A = p.pc, B = p.pc;
  if (zero_first) {
    std::fill(ref.begin(), ref.end(), Real(0));
    std::fill(B.begin(), B.end(), Real(0));  // ghosts of the destination hold zeros
    // A keeps garbage: the first launch must not read it
  }
  for (int it = 0; it < launches; ++it) {
    const bool z = zero_first && it == 0;
    double dref = 0.0, dfirst = 0.0;
    for (int m = 0; m < NIT; ++m) {
      dref = inplace_iteration<Real, Pair != 0>(p, ref);
      if (m == 0) dfirst = dref;
    }
    if constexpr (Pair != 0 && NIT == 2 && sizeof(Real) == sizeof(double)) {
      // the norm of the first iteration taken inside a two-iteration launch
      // (the reference's check iteration falling on a pass's first)
      std::vector<Real> scratch = (it & 1) ? A : B;
      const double d1 = pair_launch<Real, TI, TJ, NT, NIT, Pair == 1>(p, KS, (it & 1) ? B : A, scratch, z, order, 2);
      if (d1 != dfirst) {
        std::printf("MISMATCH first-iteration norm tile %dx%d/%d grid %dx%dx%d KS=%d launch %d: %.17g vs %.17g\n",
                    TI, TJ, NT, nx, ny, nz, KS, it + 1, d1, dfirst);
        return 1;
      }
    }
    double dfus;
    if (Pair)
      dfus = (it & 1) ? pair_launch<Real, TI, TJ, NT, NIT, Pair == 1>(p, KS, B, A, z, order)
                      : pair_launch<Real, TI, TJ, NT, NIT, Pair == 1>(p, KS, A, B, z, order);
    else
      dfus = (it & 1) ? fused_launch<Real, TI, TJ, NT, D, NIT>(p, KS, B, A, z)
                      : fused_launch<Real, TI, TJ, NT, D, NIT>(p, KS, A, B, z);
    const std::vector<Real> &cur = (it & 1) ? A : B;
    const bool norm_ok = sizeof(Real) == sizeof(double) ? dref == dfus : dfus == 0.0;
    if (std::memcmp(cur.data(), ref.data(), sizeof(Real) * ref.size()) != 0 || !norm_ok) {
      int64_t bad = -1;
      for (size_t q = 0; q < ref.size(); ++q)
        if (std::memcmp(&cur[q], &ref[q], sizeof(Real))) { bad = (int64_t)q; break; }
      std::printf("MISMATCH %s%s tile %dx%d/%d depth %d x%d grid %dx%dx%d KS=%d launch %d: first cell %lld, "
                  "delta %.17g vs %.17g\n",
                  sizeof(Real) == 8 ? "fp64" : "fp32", Pair ? (Pair == 1 ? " pair" : " pair-lds") : "", TI, TJ, NT,
                  D, NIT, nx, ny, nz, KS, it + 1,
                  (long long)bad,
                  dref, dfus);
      return 1;
    }
    if (z) std::fill(A.begin(), A.end(), Real(0));  // later launches read A: zero ghosts as the runtime keeps them
  }
  std::printf("ok %s%s tile %dx%d/%d depth %d x%d grid %dx%dx%d KS=%d launches=%d%s\n",
              sizeof(Real) == 8 ? "fp64" : "fp32", Pair ? (Pair == 1 ? " pair" : " pair-lds") : "", TI, TJ, NT, D, NIT, nx, ny, nz, KS, launches,
              zero_first ? " (zero first)" : "");
  return 0;
}

template <typename Real, int TI, int TJ, int NT, int D, int NIT = 1>
static int run_shape() {
  int bad = 0;
  bad += run_case<Real, TI, TJ, NT, D, NIT>(37, 21, 13, 4, 3, false);
  bad += run_case<Real, TI, TJ, NT, D, NIT>(TI, TJ, 9, 32, 2, false);
  bad += run_case<Real, TI, TJ, NT, D, NIT>(TI + 6, 2 * TJ + 3, 11, 3, 4, false);
  bad += run_case<Real, TI, TJ, NT, D, NIT>(5, 3, 7, 1, 3, false);
  bad += run_case<Real, TI, TJ, NT, D, NIT>(19, 9, 8, 2, 3, false);
  bad += run_case<Real, TI, TJ, NT, D, NIT>(TI + 6, 2 * TJ + 3, 11, 3, 4, true);
  return bad;
}

// the pair-owned tile (k_sor_pair)
template <typename Real, int TI, int TJ, int NT, int NIT, int CREG>
static int run_pair() {
  constexpr int P = CREG ? 1 : 2;
  int bad = 0;
  const long long steady0 = steady_steps;
  // thread orders within a step: phases in turn over all threads, then each
  // thread's whole step, ascending and descending
  for (int order = 0; order < 3; ++order) {
    // deep slabs: the middle of each march runs in guard-free groups
    bad += run_case<Real, TI, TJ, NT, 1, NIT, P>(21, 13, 40, 16, 2, false, order);
    bad += run_case<Real, TI, TJ, NT, 1, NIT, P>(37, 21, 13, 4, 3, false, order);
    bad += run_case<Real, TI, TJ, NT, 1, NIT, P>(TI, TJ, 9, 32, 2, false, order);
    bad += run_case<Real, TI, TJ, NT, 1, NIT, P>(TI + 6, 2 * TJ + 3, 11, 3, 4, false, order);
    bad += run_case<Real, TI, TJ, NT, 1, NIT, P>(5, 3, 7, 1, 3, false, order);
    bad += run_case<Real, TI, TJ, NT, 1, NIT, P>(19, 9, 8, 2, 3, false, order);
    bad += run_case<Real, TI, TJ, NT, 1, NIT, P>(TI + 6, 2 * TJ + 3, 11, 3, 4, true, order);
  }
  if (steady_steps == steady0) {
    std::printf("no guard-free step group emulated for tile %dx%d/%d x%d\n", TI, TJ, NT, NIT);
    ++bad;
  }
  return bad;
}

int main() {
  int bad = 0;
  // the FP64 shapes the library compiles (k_pcor.hip), one iteration per pass
  bad += run_shape<double, 64, 16, 256, 1>();
  bad += run_shape<double, 64, 8, 256, 1>();
  bad += run_shape<double, 64, 8, 256, 2>();
  bad += run_shape<double, 128, 16, 512, 1>();
  bad += run_shape<double, 128, 16, 512, 2>();
  bad += run_shape<double, 64, 32, 512, 1>();
  bad += run_shape<double, 64, 32, 512, 2>();
  bad += run_shape<double, 128, 8, 256, 2>();
  bad += run_shape<double, 64, 16, 512, 1>();
  bad += run_shape<double, 32, 16, 256, 1>();
  // ... and two iterations per pass
  bad += run_shape<double, 64, 16, 256, 1, 2>();
  bad += run_shape<double, 64, 16, 512, 1, 2>();
  bad += run_shape<double, 64, 32, 512, 1, 2>();
  bad += run_shape<double, 64, 8, 256, 1, 2>();
  bad += run_shape<double, 32, 16, 256, 1, 2>();
  bad += run_shape<double, 64, 32, 1024, 1>();
  bad += run_shape<double, 64, 32, 1024, 1, 2>();
  bad += run_shape<double, 128, 16, 1024, 1>();
  bad += run_shape<double, 64, 16, 1024, 1>();
  bad += run_shape<double, 128, 32, 1024, 1>();
  // the FP32 smoothing shapes (k_mixed.hip)
  bad += run_shape<float, 64, 16, 256, 1>();
  bad += run_shape<float, 64, 32, 512, 1>();
  bad += run_shape<float, 64, 32, 512, 2>();
  bad += run_shape<float, 128, 16, 512, 1>();
  bad += run_shape<float, 128, 16, 512, 2>();
  bad += run_shape<float, 128, 32, 512, 1>();
  bad += run_shape<float, 64, 16, 512, 1>();
  bad += run_shape<float, 64, 16, 256, 1, 2>();
  bad += run_shape<float, 64, 32, 512, 1, 2>();
  bad += run_shape<float, 128, 16, 512, 1, 2>();
  bad += run_shape<float, 128, 32, 512, 1, 2>();
  bad += run_shape<float, 64, 16, 512, 1, 2>();
  bad += run_shape<float, 64, 32, 1024, 1>();
  bad += run_shape<float, 64, 32, 1024, 1, 2>();
  bad += run_shape<float, 128, 32, 1024, 1>();
  bad += run_shape<float, 128, 16, 1024, 1>();
  // pair-owned tiles
  bad += run_pair<double, 64, 32, 512, 2, 1>();
  bad += run_pair<double, 64, 32, 512, 1, 1>();
  bad += run_pair<double, 64, 32, 512, 1, 0>();
  bad += run_pair<double, 64, 28, 512, 1, 0>();
  bad += run_pair<double, 64, 28, 512, 1, 1>();
  bad += run_pair<double, 64, 24, 512, 1, 0>();
  bad += run_pair<double, 64, 32, 512, 2, 0>();
  bad += run_pair<double, 64, 16, 256, 1, 1>();
  bad += run_pair<double, 64, 32, 1024, 2, 0>();
  bad += run_pair<double, 64, 32, 1024, 1, 0>();
  bad += run_pair<double, 64, 32, 1024, 2, 1>();
  bad += run_pair<double, 64, 32, 1024, 1, 1>();
  bad += run_pair<double, 128, 16, 1024, 2, 0>();
  bad += run_pair<double, 128, 16, 1024, 1, 0>();
  bad += run_pair<double, 32, 32, 512, 2, 1>();
  bad += run_pair<double, 32, 32, 512, 1, 1>();
  bad += run_pair<double, 32, 32, 512, 2, 0>();
  bad += run_pair<double, 32, 32, 512, 1, 0>();
  bad += run_pair<double, 64, 22, 512, 2, 1>();
  bad += run_pair<double, 64, 22, 512, 1, 1>();
  bad += run_pair<double, 64, 22, 512, 2, 0>();
  bad += run_pair<double, 64, 22, 512, 1, 0>();
  bad += run_pair<float, 64, 22, 1024, 2, 1>();
  bad += run_pair<float, 64, 22, 1024, 1, 1>();
  bad += run_pair<float, 64, 22, 512, 2, 1>();
  bad += run_pair<float, 64, 22, 512, 1, 1>();
  bad += run_pair<float, 64, 32, 1024, 2, 1>();
  bad += run_pair<float, 64, 32, 1024, 1, 1>();
  bad += run_pair<float, 64, 32, 512, 1, 1>();
  bad += run_pair<float, 64, 32, 512, 2, 1>();
  bad += run_pair<float, 64, 28, 512, 1, 1>();
  bad += run_pair<float, 64, 28, 512, 2, 0>();
  bad += run_pair<float, 64, 28, 512, 1, 0>();
  bad += run_pair<float, 64, 32, 512, 1, 0>();
  bad += run_pair<float, 64, 32, 512, 2, 0>();
  bad += run_pair<float, 128, 16, 512, 1, 1>();
  bad += run_pair<float, 64, 32, 1024, 2, 0>();
  bad += run_pair<float, 64, 32, 1024, 1, 0>();
  bad += run_pair<float, 64, 32, 768, 2, 1>();
  bad += run_pair<float, 64, 32, 768, 1, 1>();
  std::printf(bad ? "FAILED %d\n" : "all cases bit-identical\n", bad);
  return bad ? 1 : 0;
}
