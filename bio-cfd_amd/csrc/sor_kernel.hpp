// The fused red-black SOR iteration kernel, instantiated for the FP64 method
// (k_pcor.hip) and for the FP32 smoothing sweeps of the mixed-precision solve
// (k_mixed.hip).
#pragma once

#include "biocfd_rt.hpp"
#include "reduce.hpp"
#include "sor_fused.hpp"

namespace bcfd {

// One red-black iteration per launch on a 1-D grid of gx*gy*gz tiles; the
// tile schedule is described with SorTile.  Non-fluid cells of the tile are
// copied through, so src and dst agree on every cell the iteration leaves.
// On check iterations (every fifth, numbered by the host) each workgroup
// posts its largest update to partials.
// MINB: workgroups per CU the register allocation must leave room for
template <typename Real, int TI, int TJ, int NT, int D, int MINB>
__global__ void __launch_bounds__(NT, MINB) k_sor_fused(FusedArgsT<Real> a, int gx, int gy, int check,
                                                  const SorState *__restrict__ st, double *__restrict__ partials) {
  using T = SorTile<Real, TI, TJ, NT, D>;
  if (st->done) return;
  __shared__ Real lds[T::LDS_DOUBLES];
  const int tid = threadIdx.x;
  const typename T::Tile tile = T::tile_of(a, T::xcd_tile(blockIdx.x, gridDim.x), gx, gy);
  typename T::Thread th;
  T::init_thread(a, tile, tid, th);
  T::load_coef(a, tile, lds, tid);
  T::prologue(a, tile, lds, th);
  double delta = 0.0;
  for (int s = tile.k0 - 1; s <= tile.k1 + 1; s += D) {
    __syncthreads();
    T::phase_a(a, tile, lds, th, s, th.pcn[0]);
    __syncthreads();
    const double dr = T::phase_b(a, lds, th, s);
    delta = delta > dr ? delta : dr;
    __syncthreads();
    const double db = T::phase_c(a, tile, lds, th, s);
    delta = delta > db ? delta : db;
    if (D > 1 && s + 1 <= tile.k1 + 1) {
      __syncthreads();
      T::phase_a(a, tile, lds, th, s + 1, th.pcn[D - 1]);
      __syncthreads();
      const double dr2 = T::phase_b(a, lds, th, s + 1);
      delta = delta > dr2 ? delta : dr2;
      __syncthreads();
      const double db2 = T::phase_c(a, tile, lds, th, s + 1);
      delta = delta > db2 ? delta : db2;
    }
  }
  if (check) wg_max_store(delta, partials);
}

}  // namespace bcfd
