// The fused red-black SOR iteration kernel, instantiated for the FP64 method
// (k_pcor.hip) and for the FP32 smoothing sweeps of the mixed-precision solve
// (k_mixed.hip).
#pragma once

#include "biocfd_rt.hpp"
#include "reduce.hpp"
#include "sor_fused.hpp"
#include "sor_pair.hpp"

namespace bcfd {

// NIT red-black iterations per launch on a 1-D grid of gx*gy*gz tiles; the
// tile schedule is described with SorTile.  Non-fluid cells of the tile are
// copied through, so src and dst agree on every cell the launch leaves.
// When `check` is set (the launch's last iteration is a check iteration,
// every fifth, numbered by the host) each workgroup posts the largest update
// of that iteration to partials.
// MINB: workgroups per CU the register allocation must leave room for
template <typename Real, int TI, int TJ, int NT, int D, int NIT, int MINB>
__global__ void __launch_bounds__(NT, MINB) k_sor_fused(FusedArgsT<Real> a, int gx, int gy, int check,
                                                        const SorState *__restrict__ st,
                                                        double *__restrict__ partials) {
  using T = SorTile<Real, TI, TJ, NT, D, NIT>;
  if (st->done) return;
  __shared__ Real lds[T::LDS_DOUBLES];
  const int tid = threadIdx.x;
  const typename T::Tile tile = T::tile_of(a, T::xcd_tile(blockIdx.x, gridDim.x), gx, gy);
  typename T::Thread th;
  T::init_thread(a, tile, tid, th);
  T::load_coef(a, tile, lds, tid);
  T::prologue(a, tile, lds, th);
  double delta = 0.0;
  const int s1 = T::last_step(tile);
  for (int s = T::first_step(tile); s <= s1; s += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (d > 0 && s + d > s1) break;
      __syncthreads();
      T::phase_a(a, tile, lds, th, s + d, d);
#pragma unroll
      for (int p = 0; p < T::NPH; ++p) {
        __syncthreads();
        const double dp = T::phase_u(a, tile, lds, th, p, s + d, check && p >= T::NPH - 2);
        delta = delta > dp ? delta : dp;
      }
    }
  }
  if (check) wg_max_store(delta, partials);
}

// one step of the pair-owned tile march; B: the step's rhs buffer, S: a
// step of a guard-free group.  One barrier: within a step no thread reads an
// LDS word another thread writes (SorPairTile).  check != 0: the largest
// update of the pass's last iteration (its last two phases), or with FIRST of
// its first iteration (phases 0 and 1; the cells a phase updates outside the
// tile are cells of the block computed as their own tile computes them, so
// the maximum over all workgroups is the iteration's).  FIRST is a template
// parameter so that each kernel carries the norm code of two phases only.
template <class T, int B, bool S, bool FIRST, typename Real>
__device__ __forceinline__ void pair_step(const FusedArgsT<Real> &a, const typename T::Tile &tile, Real *lds,
                                 typename T::Thread &th, int s, int check, double &delta) {
  __syncthreads();
  T::template step_begin<B, S>(a, tile, lds, th, s);
#pragma unroll
  for (int p = 0; p < T::NPH; ++p) {
    const bool norm = check && (FIRST ? p < 2 : p >= T::NPH - 2);
    const double dp = T::template phase_u<B, S>(a, tile, lds, th, p, s, norm);
    delta = delta > dp ? delta : dp;
  }
}

// RB steps from s (the last ones only while s + b <= s1 unless S)
template <class T, bool S, bool FIRST, typename Real>
__device__ __forceinline__ void pair_group(const FusedArgsT<Real> &a, const typename T::Tile &tile, Real *lds,
                                  typename T::Thread &th, int s, int s1, int check, double &delta) {
  pair_step<T, 0, S, FIRST>(a, tile, lds, th, s, check, delta);
  if (S || s + 1 <= s1) pair_step<T, 1, S, FIRST>(a, tile, lds, th, s + 1, check, delta);
  if constexpr (T::RB == 4) {
    if (S || s + 2 <= s1) pair_step<T, 2 % T::RB, S, FIRST>(a, tile, lds, th, s + 2, check, delta);
    if (S || s + 3 <= s1) pair_step<T, 3 % T::RB, S, FIRST>(a, tile, lds, th, s + 3, check, delta);
  }
}

// the pair-owned tile march of one workgroup on tile `tile_index` (1-D,
// gx*gy tiles per z chunk) of the block described by `a`
template <class T, typename Real, bool FIRST = false>
__device__ __forceinline__ void pair_run(const FusedArgsT<Real> &a, int tile_index, int gx, int gy, int check,
                                double *__restrict__ partials) {
  __shared__ Real lds[T::LDS_DOUBLES];
  const typename T::Tile tile = T::tile_of(a, tile_index, gx, gy);
  typename T::Thread th;
  T::init_thread(a, tile, threadIdx.x, th);
  T::prologue(a, tile, lds, th, threadIdx.x);
  double delta = 0.0;
  const int s1 = T::last_step(tile);
  // groups of RB steps, guard-free where the whole group fits the range
  // that needs no guard
  for (int s = T::first_step(tile); s <= s1; s += T::RB) {
    if (T::group_steady(a, tile, s))
      pair_group<T, true, FIRST>(a, tile, lds, th, s, s1, check, delta);
    else
      pair_group<T, false, FIRST>(a, tile, lds, th, s, s1, check, delta);
  }
  if (check) wg_max_store(delta, partials);
}

// NIT red-black iterations per launch with the pair-owned tile (SorPairTile);
// same grid, arguments and check protocol as k_sor_fused.  MINW: waves per
// SIMD the register allocation must leave room for (4 = two 512-thread
// workgroups per CU)
// FIRST: a check pass takes the norm of its first iteration (the reference's
// check iteration falling on a two-iteration pass's first)
template <typename Real, int TI, int TJ, int NT, int NIT, int MINW, bool CREG, bool FIRST = false>
__global__ void __launch_bounds__(NT, MINW) k_sor_pair(FusedArgsT<Real> a, int gx, int gy, int check,
                                                       const SorState *__restrict__ st,
                                                       double *__restrict__ partials) {
  using T = SorPairTile<Real, TI, TJ, NT, NIT, CREG>;
  if (st->done) return;
  pair_run<T, Real, FIRST>(a, T::xcd_tile(blockIdx.x, gridDim.x), gx, gy, check, partials);
}

// The rollback pass of a call that stopped on the first iteration of a
// two-iteration pass: that iteration alone (NIT = 1) from the pass's source,
// the launch with parity = the number of passes before it (mod 2), so src /
// dst are that pass's; every other launch returns at once.  Kept apart from
// k_sor_pair so that the iteration kernels carry none of its logic.
template <typename Real, int TI, int TJ, int NT, int MINW, bool CREG>
__global__ void __launch_bounds__(NT, MINW) k_sor_pair_rollback(FusedArgsT<Real> a, int gx, int gy, int parity,
                                                                const SorState *__restrict__ st) {
  using T = SorPairTile<Real, TI, TJ, NT, 1, CREG>;
  if (!st->rollback || (st->passes & 1) != parity) return;
  pair_run<T, Real>(a, T::xcd_tile(blockIdx.x, gridDim.x), gx, gy, 0, nullptr);
}

// One block of a batched launch: its arguments for passes that read pc
// (parity 0) and pc2 (parity 1), its tile grid, the first workgroup of the
// launch that works on it, and its SOR state.
template <typename Real>
struct BatchBlockT {
  FusedArgsT<Real> a[2];
  int gx, gy, ntiles, first;
  SorState *st;
  long long start;   // nIterPcor at the start of the block's call
};

// The same iterations on every block of a set in one launch: workgroups
// first .. first + ntiles - 1 work on block b of the table.  The blocks'
// partials follow each other (workgroup index of the launch); a block whose
// SOR call has stopped skips its tiles.
template <typename Real, int TI, int TJ, int NT, int NIT, int MINW, bool CREG, bool FIRST = false>
__global__ void __launch_bounds__(NT, MINW) k_sor_pair_batch(const BatchBlockT<Real> *__restrict__ tab, int nblk,
                                                             int parity, int check, double *__restrict__ partials) {
  using T = SorPairTile<Real, TI, TJ, NT, NIT, CREG>;
  int b = 0;
  while (b + 1 < nblk && (int)blockIdx.x >= tab[b + 1].first) ++b;
  if (tab[b].st->done) return;
  const FusedArgsT<Real> a = tab[b].a[parity];
  pair_run<T, Real, FIRST>(a, T::xcd_tile((int)blockIdx.x - tab[b].first, tab[b].ntiles), tab[b].gx, tab[b].gy, check,
                    partials);
}

// the rollback pass of every block of a batched set that stopped on a pass's
// first iteration (see k_sor_pair_rollback)
template <typename Real, int TI, int TJ, int NT, int MINW, bool CREG>
__global__ void __launch_bounds__(NT, MINW) k_sor_pair_batch_rollback(const BatchBlockT<Real> *__restrict__ tab,
                                                                      int nblk) {
  using T = SorPairTile<Real, TI, TJ, NT, 1, CREG>;
  int b = 0;
  while (b + 1 < nblk && (int)blockIdx.x >= tab[b + 1].first) ++b;
  const SorState *st = tab[b].st;
  if (!st->rollback) return;
  const FusedArgsT<Real> a = tab[b].a[st->passes & 1];
  pair_run<T, Real>(a, T::xcd_tile((int)blockIdx.x - tab[b].first, tab[b].ntiles), tab[b].gx, tab[b].gy, 0, nullptr);
}

}  // namespace bcfd
