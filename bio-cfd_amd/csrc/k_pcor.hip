// Pressure-correction projection on gfx950: divergence, red-black SOR,
// pressure/velocity correction, domain boundary conditions, steady-state
// norm and the poissonSolver orchestration (reference src/PcorVcor.F90 and
// src/BoundaryConditions.f90).
//
// FP64 and bit-identical to the reference CPU build: the library is built
// with -ffp-contract=off and every expression keeps the reference's operand
// order.  Kernels sweep the dense grid with the cell tag as the fluid mask
// (fluid == tag 0 at interior cells), which visits exactly the cells of the
// reference's fluid/red/black index lists; no per-cell index list is read.
//
// SOR storage: a single pc array.  The reference keeps pco == pc at the start
// of every iteration (PcorVcor.F90:363-372), red cells read only black
// neighbours and black cells read the new red values, so updating pc in place
// colour by colour is the same computation; |pc - pco| of an iteration is
// the per-cell change, reduced with one atomic per wavefront.
#include <cstring>

#include "biocfd_rt.hpp"
#include "reduce.hpp"
#include "sor_kernel.hpp"

namespace bcfd {

int interface_coarse_update(int which);
int interface_fine_update(int which);


__global__ void k_zero_pcor(double *__restrict__ b, double *__restrict__ pc, int64_t n) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) { b[q] = 0.0; pc[q] = 0.0; }
}

__global__ void k_sor_reset_call(SorState *st, int reset_iter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (reset_iter) { st->n_iter = 0; st->derr2 = 0.0; st->derr_std_st = 0.0; st->exhausted = 0; }
    st->derr4 = 11111.0;
    st->max_bits = 0ull;
    st->done = 0;
    st->call_iters = 0;
    st->passes = 0;
    st->rollback = 0;
  }
}

// computeDiv (PcorVcor.F90:222-243) over fluid cells
__global__ void k_div(Dims d, const uint8_t *__restrict__ cell, const double *__restrict__ ut,
                      const double *__restrict__ vt, const double *__restrict__ wt,
                      const double *__restrict__ dxv, const double *__restrict__ dyv,
                      const double *__restrict__ dzv, double *__restrict__ b) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (i > d.nx + 1 || j > d.ny + 1) return;
  const int64_t c = ix3(d, i, j, k);
  if (cell[c] != 0) return;
  b[c] = (ut[c] - ut[c - 1]) / dxv[i - 1] + (vt[c] - vt[c - d.NX]) / dyv[j - 1] +
         (wt[c] - wt[c - (int64_t)d.NX * d.NY]) / dzv[k - 1];
}

// Colour-split storage of pc and b for the SOR iterations: colour 1 (red,
// (i+j+k) odd) and colour 0 (black) each hold their cells of every row in
// a contiguous half-row, h = (i-1)>>1.  In row (j,k) the cells of colour c
// sit at i = 2h + 2 - ((c - j - k) & 1).  A colour sweep then reads its own
// half-array, its b half-array and the other colour's half-array once, fully
// coalesced.

__global__ void k_split(Dims d, int NXh, const double *__restrict__ full, double *__restrict__ red,
                        double *__restrict__ black) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.NX || j > d.NY) return;
  const double v = full[ix3(d, i, j, k)];
  double *dst = ((i + j + k) & 1) ? red : black;
  dst[sidx(NXh, d.NY, (i - 1) >> 1, j, k)] = v;
}

__global__ void k_merge(Dims d, int NXh, double *__restrict__ full, const double *__restrict__ red,
                        const double *__restrict__ black) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.NX || j > d.NY) return;
  const double *src = ((i + j + k) & 1) ? red : black;
  full[ix3(d, i, j, k)] = src[sidx(NXh, d.NY, (i - 1) >> 1, j, k)];
}


// b/dt of the fluid cells in the split layout; non-fluid cells carry a NaN
// with a private payload so the sweeps need no tag array (b/dt is the same
// division the reference performs inside every sweep, done once)

__global__ void k_split_bdt(Dims d, int NXh, const double *__restrict__ b, const uint8_t *__restrict__ cell,
                            double dt, double *__restrict__ red, double *__restrict__ black) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > d.NX || j > d.NY) return;
  const int64_t c = ix3(d, i, j, k);
  const bool interior = i >= 2 && i <= d.nx + 1 && j >= 2 && j <= d.ny + 1 && k >= 2 && k <= d.nz + 1;
  const double v = (interior && cell[c] == 0) ? b[c] / dt : __longlong_as_double((long long)kNotFluid);
  double *dst = ((i + j + k) & 1) ? red : black;
  dst[sidx(NXh, d.NY, (i - 1) >> 1, j, k)] = v;
}

constexpr int kSorKz = 4;  // planes per thread

// One colour of REDBLACKSOR_linear (PcorVcor.F90:313-330 red / 332-349 black)
// on the split layout; `other` holds the neighbours' colour.  Thread t of a
// row owns the t-th cell of the colour (i = i0 + 2t) on kSorKz planes.
__global__ void __launch_bounds__(256) k_sor_split(Dims d, int NXh, int colour, const double *__restrict__ bdt,
                                                   double *__restrict__ own, const double *__restrict__ other,
                                                   const double *__restrict__ Ax, const double *__restrict__ Ay,
                                                   const double *__restrict__ Az, double omega,
                                                   const SorState *__restrict__ st, double *__restrict__ partials) {
  if (st->done) return;
  const bool check = ((st->n_iter + 1) % 5) == 0;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = d.nx, ny = d.ny, nz = d.nz, NY = d.NY;
  const int64_t sy = NXh, sz = (int64_t)NXh * NY;
  double delta = 0.0;
  if (j <= ny + 1) {
    const double axy = Ay[(j - 2) + ny];
    const double ays = Ay[(j - 2)], ayn = Ay[(j - 2) + 2 * ny];
    for (int kk = 0; kk < kSorKz; ++kk) {
      const int k = 2 + blockIdx.z * kSorKz + kk;
      if (k > nz + 1) break;
      const int i = 2 + ((colour - j - k) & 1) + 2 * t;
      if (i > nx + 1) continue;
      const int h = (i - 1) >> 1;
      const int64_t c = sidx(NXh, NY, h, j, k);
      const double rhs = bdt[c];
      if (__double_as_longlong(rhs) == (long long)kNotFluid) continue;
      const int64_t cw = sidx(NXh, NY, (i - 2) >> 1, j, k), ce = sidx(NXh, NY, i >> 1, j, k);
      const double old = own[c];
      const double gs = (rhs - Ax[(i - 2)] * other[cw] - Ax[(i - 2) + 2 * nx] * other[ce] - ays * other[c - sy] -
                         ayn * other[c + sy] - Az[(k - 2)] * other[c - sz] - Az[(k - 2) + 2 * nz] * other[c + sz]) /
                        (Ax[(i - 2) + nx] + axy + Az[(k - 2) + nz]);
      const double nw = (1.0 - omega) * old + omega * gs;
      own[c] = nw;
      const double dd = fabs(nw - old);
      delta = delta > dd ? delta : dd;
    }
  }
  if (check) wg_max_store(delta, partials);
}

// Bookkeeping of fused iteration `it` (PcorVcor.F90:309, 351-375), launched
// after check iterations and after the last iteration the call may run: the
// norm only changes on check iterations and the pcItaMax exit only fires on
// the last one, so the state after any other iteration needs no update until
// the next of these.
__global__ void __launch_bounds__(1024) k_sor_fused_end(SorState *st, long long it, long long start,
                                                        long long passes, int check, long long pc_ita_max,
                                                        double epsi, const double *partials, int64_t n_partials,
                                                        int first = 0) {
  if (st->done) return;
  __shared__ double s_m[16];
  if (check) {
    double m = 0.0;
    for (int64_t q = threadIdx.x; q < n_partials; q += blockDim.x) m = m > partials[q] ? m : partials[q];
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (check) {
    double m = 0.0;
    for (int q = 0; q < (int)(blockDim.x + 63) / 64; ++q) m = m > s_m[q] ? m : s_m[q];
    st->derr4 = m;
  }
  if (first && check && st->derr4 < epsi) {
    // stopped on the pass's first iteration (it - 1): that iteration is
    // redone alone from the pass's source, which counts as the pass
    st->n_iter = it - 1;
    st->call_iters = it - 1 - start;
    st->passes = (int)passes - 1;
    st->derr2 = st->derr4;
    st->done = 1;
    st->rollback = 1;
    return;
  }
  st->n_iter = it;
  st->call_iters = it - start;
  st->passes = (int)passes;
  st->derr2 = st->derr4;
  if (st->derr4 < epsi || it > pc_ita_max) {
    st->done = 1;
    st->exhausted += !(st->derr4 < epsi);
  }
}

// k_sor_fused_end with the iteration and pass counts relative to the
// previous end kernel of the call (dit iterations and dpass passes later),
// which set n_iter, call_iters and passes: the same state updates, with
// arguments that do not depend on where the call started, so that a
// captured chunk of passes can be replayed
__global__ void __launch_bounds__(1024) k_sor_fused_end_rel(SorState *st, long long dit, long long dpass, int check,
                                                            long long pc_ita_max, double epsi,
                                                            const double *partials, int64_t n_partials,
                                                            int first) {
  if (st->done) return;
  __shared__ double s_m[16];
  if (check) {
    double m = 0.0;
    for (int64_t q = threadIdx.x; q < n_partials; q += blockDim.x) m = m > partials[q] ? m : partials[q];
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (check) {
    double m = 0.0;
    for (int q = 0; q < (int)(blockDim.x + 63) / 64; ++q) m = m > s_m[q] ? m : s_m[q];
    st->derr4 = m;
  }
  const long long it = st->n_iter + dit;
  if (first && check && st->derr4 < epsi) {
    st->n_iter = it - 1;
    st->call_iters += dit - 1;
    st->passes += (int)dpass - 1;
    st->derr2 = st->derr4;
    st->done = 1;
    st->rollback = 1;
    return;
  }
  st->n_iter = it;
  st->call_iters += dit;
  st->passes += (int)dpass;
  st->derr2 = st->derr4;
  if (st->derr4 < epsi || it > pc_ita_max) {
    st->done = 1;
    st->exhausted += !(st->derr4 < epsi);
  }
}

// after the rollback pass: it counts as the stopping pass
__global__ void k_sor_rollback_end(SorState *st) {
  if (threadIdx.x == 0 && st->rollback) {
    st->passes += 1;
    st->rollback = 0;
  }
}
__global__ void k_sor_rollback_end_set(const BatchBlockT<double> *__restrict__ tab, int nblk) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  SorState *st = tab[b].st;
  if (st->rollback) {
    st->passes += 1;
    st->rollback = 0;
  }
}

// k_sor_fused_end for every block of a batched launch: workgroup b reduces
// block b's partials (workgroups first .. first + ntiles - 1 of the launch);
// `done` iterations of the set have run, so block b is at iteration
// start_b + done
__global__ void __launch_bounds__(1024) k_sor_batch_end(const BatchBlockT<double> *__restrict__ tab,
                                                        long long done, long long passes, int check,
                                                        long long pc_ita_max, double epsi,
                                                        const double *__restrict__ partials, int first) {
  const BatchBlockT<double> &e = tab[blockIdx.x];
  SorState *st = e.st;
  if (st->done) return;
  __shared__ double s_m[16];
  if (check) {
    double m = 0.0;
    for (int64_t q = threadIdx.x; q < e.ntiles; q += blockDim.x) {
      const double v = partials[e.first + q];
      m = m > v ? m : v;
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (check) {
    double m = 0.0;
    for (int q = 0; q < (int)(blockDim.x + 63) / 64; ++q) m = m > s_m[q] ? m : s_m[q];
    st->derr4 = m;
  }
  const long long it = e.start + done;
  if (first && check && st->derr4 < epsi) {
    st->n_iter = it - 1;
    st->call_iters = done - 1;
    st->passes = (int)passes - 1;
    st->derr2 = st->derr4;
    st->done = 1;
    st->rollback = 1;
    return;
  }
  st->n_iter = it;
  st->call_iters = done;
  st->passes = (int)passes;
  st->derr2 = st->derr4;
  if (st->derr4 < epsi || it > pc_ita_max) {
    st->done = 1;
    st->exhausted += !(st->derr4 < epsi);
  }
}

// k_sor_batch_end with counts relative to the previous end kernel of the
// set's calls (dit iterations and dpass passes later), for captured chunks
__global__ void __launch_bounds__(1024) k_sor_batch_end_rel(const BatchBlockT<double> *__restrict__ tab,
                                                            long long dit, long long dpass, int check,
                                                            long long pc_ita_max, double epsi,
                                                            const double *__restrict__ partials, int first) {
  const BatchBlockT<double> &e = tab[blockIdx.x];
  SorState *st = e.st;
  if (st->done) return;
  __shared__ double s_m[16];
  if (check) {
    double m = 0.0;
    for (int64_t q = threadIdx.x; q < e.ntiles; q += blockDim.x) {
      const double v = partials[e.first + q];
      m = m > v ? m : v;
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (check) {
    double m = 0.0;
    for (int q = 0; q < (int)(blockDim.x + 63) / 64; ++q) m = m > s_m[q] ? m : s_m[q];
    st->derr4 = m;
  }
  const long long it = st->n_iter + dit;
  if (first && check && st->derr4 < epsi) {
    st->n_iter = it - 1;
    st->call_iters += dit - 1;
    st->passes += (int)dpass - 1;
    st->derr2 = st->derr4;
    st->done = 1;
    st->rollback = 1;
    return;
  }
  st->n_iter = it;
  st->call_iters += dit;
  st->passes += (int)dpass;
  st->derr2 = st->derr4;
  if (st->derr4 < epsi || it > pc_ita_max) {
    st->done = 1;
    st->exhausted += !(st->derr4 < epsi);
  }
}

// the passes of one SOR call alternate pc -> pc2 -> pc ...; after an odd
// count the result is in pc2
__global__ void k_sor_finish(const SorState *__restrict__ st, double *__restrict__ pc,
                             const double *__restrict__ pc2, int64_t n) {
  if ((st->passes & 1) == 0) return;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    pc[q] = pc2[q];
}

// end-of-iteration bookkeeping of REDBLACKSOR_linear (PcorVcor.F90:309, 351-375);
// on check iterations reduces the sweeps' workgroup maxima (one workgroup)
__global__ void __launch_bounds__(1024) k_sor_iter_end(SorState *st, long long pc_ita_max, double epsi,
                                                       const double *partials, int64_t n_partials) {
  if (st->done) return;
  const long long it = st->n_iter + 1;
  __shared__ double s_m[16];
  if (it % 5 == 0) {
    double m = 0.0;
    for (int64_t q = threadIdx.x; q < n_partials; q += blockDim.x) m = m > partials[q] ? m : partials[q];
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  st->n_iter = it;
  st->call_iters += 1;
  if (it % 5 == 0) {
    double m = 0.0;
    for (int q = 0; q < (int)(blockDim.x + 63) / 64; ++q) m = m > s_m[q] ? m : s_m[q];
    st->derr4 = m;
  }
  st->derr2 = st->derr4;
  if (st->derr4 < epsi || it > pc_ita_max) {
    st->done = 1;
    st->exhausted += !(st->derr4 < epsi);
  }
}

// correctPressure + correctVelocity (PcorVcor.F90:245-286)
__global__ void k_correct(Dims d, const uint8_t *__restrict__ cell, const double *__restrict__ pc,
                          double *__restrict__ p, double *__restrict__ ut, double *__restrict__ vt,
                          double *__restrict__ wt, const double *__restrict__ dxv,
                          const double *__restrict__ dyv, const double *__restrict__ dzv, double dt) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (i > d.nx + 1 || j > d.ny + 1) return;
  const int64_t c = ix3(d, i, j, k);
  if (cell[c] != 0) return;
  const int64_t sy = d.NX, sz = (int64_t)d.NX * d.NY;
  const double pcc = pc[c];
  p[c] = p[c] + pcc;
  ut[c] = ut[c] - dt / (0.5 * (dxv[i] + dxv[i - 1])) * (pc[c + 1] - pcc);
  vt[c] = vt[c] - dt / (0.5 * (dyv[j] + dyv[j - 1])) * (pc[c + sy] - pcc);
  wt[c] = wt[c] - dt / (0.5 * (dzv[k] + dzv[k - 1])) * (pc[c + sz] - pcc);
}

// steady-state norm over fluid cells, then u,v,w <- ut,vt,wt everywhere
// (PcorVcor.F90:135-150, 187-199)
__global__ void k_norm_copy(Dims d, const uint8_t *__restrict__ cell, double *__restrict__ u,
                            double *__restrict__ v, double *__restrict__ w,
                            const double *__restrict__ ut, const double *__restrict__ vt,
                            const double *__restrict__ wt, double dt, double *__restrict__ partials) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  double m = 0.0;
  if (i <= d.NX && j <= d.NY) {
    const int64_t c = ix3(d, i, j, k);
    const double a = ut[c], b = vt[c], e = wt[c];
    if (i >= 2 && i <= d.nx + 1 && j >= 2 && j <= d.ny + 1 && k >= 2 && k <= d.nz + 1 && cell[c] == 0) {
      const double eu = fabs((a - u[c])) / dt;
      const double ev = fabs((b - v[c])) / dt;
      const double ew = fabs((e - w[c])) / dt;
      m = eu > ev ? eu : ev;
      m = m > ew ? m : ew;
    }
    u[c] = a; v[c] = b; w[c] = e;
  }
  wg_max_store(m, partials);
}

__global__ void __launch_bounds__(1024) k_norm_end(SorState *st, const double *partials, int64_t n) {
  __shared__ double s_m[16];
  double m = 0.0;
  for (int64_t q = threadIdx.x; q < n; q += blockDim.x) m = m > partials[q] ? m : partials[q];
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int q = 0; q < (int)(blockDim.x + 63) / 64; ++q) r = r > s_m[q] ? r : s_m[q];
    st->derr_std_st = r;
  }
}

// velocityBC (BoundaryConditions.f90:11-74): three face passes in order.
// `inlet` is the inflow value (uc; 0 for the tangent-linear pass, where uc
// stays the convective speed of the outflow condition)
__global__ void k_bc_x(Dims d, const double *__restrict__ u, const double *__restrict__ v,
                       const double *__restrict__ w, double *__restrict__ ut, double *__restrict__ vt,
                       double *__restrict__ wt, const double *__restrict__ dxv, double dt, double uc,
                       double inlet) {
  const int j = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int k = 2 + blockIdx.y;
  if (j > d.ny + 1) return;
  const int nx = d.nx;
  const int64_t c1 = ix3(d, 1, j, k), cn = ix3(d, nx, j, k), cn1 = ix3(d, nx + 1, j, k),
                cn2 = ix3(d, nx + 2, j, k);
  ut[c1] = inlet;
  vt[c1] = -vt[cn1];
  wt[c1] = -wt[cn1];
  ut[cn1] = u[cn1] - (dt / dxv[nx]) * uc * (u[cn1] - u[cn]);
  vt[cn2] = -vt[cn1] + v[cn2] + v[cn1] - (2.0 * dt / dxv[nx + 1]) * uc * (v[cn2] - v[cn1]);
  wt[cn2] = -wt[cn1] + w[cn2] + w[cn1] - (2.0 * dt / dxv[nx + 1]) * uc * (w[cn2] - w[cn1]);
}

__global__ void k_bc_y(Dims d, double *__restrict__ ut, double *__restrict__ vt, double *__restrict__ wt) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int k = 2 + blockIdx.y;
  if (i > d.nx + 1) return;
  const int ny = d.ny;
  ut[ix3(d, i, 1, k)] = ut[ix3(d, i, 2, k)];
  wt[ix3(d, i, 1, k)] = wt[ix3(d, i, 2, k)];
  vt[ix3(d, i, 2, k)] = 0.0;
  ut[ix3(d, i, ny + 2, k)] = ut[ix3(d, i, ny + 1, k)];
  wt[ix3(d, i, ny + 2, k)] = wt[ix3(d, i, ny + 1, k)];
  vt[ix3(d, i, ny + 1, k)] = 0.0;
}

__global__ void k_bc_z(Dims d, double *__restrict__ ut, double *__restrict__ vt, double *__restrict__ wt) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y;
  if (i > d.nx + 1) return;
  const int nz = d.nz;
  ut[ix3(d, i, j, 1)] = ut[ix3(d, i, j, 2)];
  vt[ix3(d, i, j, 1)] = vt[ix3(d, i, j, 2)];
  wt[ix3(d, i, j, 1)] = 0.0;
  ut[ix3(d, i, j, nz + 2)] = ut[ix3(d, i, j, nz + 1)];
  vt[ix3(d, i, j, nz + 2)] = vt[ix3(d, i, j, nz + 1)];
  wt[ix3(d, i, j, nz + 1)] = 0.0;
}

// solidCellBC / solidCellBC_move: zero the listed solid cells (tag 1)
__global__ void k_solid_bc(Dims d, const uint8_t *__restrict__ cell, double *__restrict__ ut,
                           double *__restrict__ vt, double *__restrict__ wt, double *__restrict__ p,
                           double *__restrict__ u, double *__restrict__ v, double *__restrict__ w,
                           int also_u) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (i > d.nx + 1 || j > d.ny + 1) return;
  const int64_t c = ix3(d, i, j, k);
  if (cell[c] != 1) return;
  ut[c] = 0.0; vt[c] = 0.0; wt[c] = 0.0; p[c] = 0.0;
  if (also_u) { u[c] = 0.0; v[c] = 0.0; w[c] = 0.0; }
}

__global__ void k_copy3(double *__restrict__ a, double *__restrict__ b, double *__restrict__ c,
                        const double *__restrict__ x, const double *__restrict__ y,
                        const double *__restrict__ z, int64_t n) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) { a[q] = x[q]; b[q] = y[q]; c[q] = z[q]; }
}

static dim3 interior_grid(const Dims &d, dim3 blk) {
  return dim3(div_up(d.nx, blk.x), div_up(d.ny, blk.y), (unsigned)d.nz);
}

int poisson_block_setup(Block &b) {
  (void)b;
  return 0;
}

static int check_ready(const Block &b) {
  for (int q : {BG_DX, BG_DY, BG_DZ})
    if (!b.g[q]) { set_error("block grids not uploaded"); return -1; }
  if (!b.A[0]) { set_error("coefficientMatrix has not been called"); return -1; }
  return 0;
}

int compute_div(Block &b) {
  hipStream_t s = rt().stream;
  dim3 blk(64, 4, 1);
  hipLaunchKernelGGL(k_div, interior_grid(b.d, blk), blk, 0, s, b.d, b.tag[BT_CELL], b.f[BF_UT],
                     b.f[BF_VT], b.f[BF_WT], b.g[BG_DX], b.g[BG_DY], b.g[BG_DZ], b.f[BF_B]);
  BCFD_CHECK_LAUNCH();
  return 0;
}

// REDBLACKSOR_linear: iterations are enqueued in chunks; the device decides
// when to stop (done flag), the host only polls between chunks.
// optional HIP-event timing of the SOR sweeps (events on the runtime stream)
struct SorTiming {
  bool on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  int64_t iters = 0;
  int64_t cells = 0;  // fluid cells swept per iteration, summed over calls
};
static SorTiming g_tm;

int ensure_split(Block &b) {
  const int NXh = (b.d.NX + 1) / 2;
  const int64_t n = (int64_t)NXh * b.d.NY * b.d.NZ;
  if (b.split_n != n) {
    for (int c = 0; c < 2; ++c)
      if (dev_alloc(&b.pcs[c], n) || dev_alloc(&b.bs[c], n)) return -1;
    b.split_n = n;
  }
  // one partial per workgroup of each colour sweep (the norm copy needs fewer)
  const int64_t nwg = (int64_t)div_up(NXh, 64) * div_up(b.d.ny, 4) * b.d.nz;
  if (grow(&b.partials, &b.cap_partials, 2 * nwg + (int64_t)div_up(b.d.NX, 64) * div_up(b.d.NY, 4) * b.d.NZ))
    return -1;
  return 0;
}

// SOR kernel variants: 0 = one launch per colour on the colour-split arrays
// (plus one for the norm), 1.. = fused launches of `nit` iterations with the
// tile shapes and prefetch depths below (a pass shortened by a check
// iteration or by the budget runs the same tile with one iteration)
using FusedLaunch = void (*)(dim3, dim3, hipStream_t, const FusedArgs &, int, int, int, const SorState *, double *);
using BatchBlock = BatchBlockT<double>;
using BatchLaunch = void (*)(dim3, dim3, hipStream_t, const BatchBlock *, int, int, int, double *);
using RollbackLaunch = void (*)(dim3, dim3, hipStream_t, const FusedArgs &, int, int, int, const SorState *);
using BatchRollbackLaunch = void (*)(dim3, dim3, hipStream_t, const BatchBlock *, int);
struct FusedVariant {
  int ti, tj, nt, depth, minb, nit;
  FusedLaunch launch, launch1;
  int family;  // 0: SorTile (cell lists per phase), 1, 2: SorPairTile (pair-owned; coefficients in registers, LDS)
  BatchLaunch batch, batch1;  // several blocks in one launch (pair-owned tiles only)
  // two-iteration pair-owned passes whose check iteration is their first,
  // and the rollback pass after a stop on such an iteration
  FusedLaunch launch_first;
  BatchLaunch batch_first;
  RollbackLaunch rollback;
  BatchRollbackLaunch batch_rollback;
};

template <int TI, int TJ, int NT, int D, int MINB, int NIT>
static void launch_fused(dim3 g, dim3 blk, hipStream_t s, const FusedArgs &a, int gx, int gy, int check,
                         const SorState *st, double *partials) {
  hipLaunchKernelGGL((k_sor_fused<double, TI, TJ, NT, D, NIT, MINB>), g, blk, 0, s, a, gx, gy, check, st,
                     partials);
}

template <int TI, int TJ, int NT, int MINW, int NIT, bool CREG, bool FIRST = false>
static void launch_pair(dim3 g, dim3 blk, hipStream_t s, const FusedArgs &a, int gx, int gy, int check,
                        const SorState *st, double *partials) {
  hipLaunchKernelGGL((k_sor_pair<double, TI, TJ, NT, NIT, MINW, CREG, FIRST>), g, blk, 0, s, a, gx, gy, check, st,
                     partials);
}

template <int TI, int TJ, int NT, int MINW, int NIT, bool CREG, bool FIRST = false>
static void launch_pair_batch(dim3 g, dim3 blk, hipStream_t s, const BatchBlock *tab, int nblk, int parity, int check,
                              double *partials) {
  hipLaunchKernelGGL((k_sor_pair_batch<double, TI, TJ, NT, NIT, MINW, CREG, FIRST>), g, blk, 0, s, tab, nblk, parity,
                     check, partials);
}

template <int TI, int TJ, int NT, int MINW, bool CREG>
static void launch_pair_rollback(dim3 g, dim3 blk, hipStream_t s, const FusedArgs &a, int gx, int gy, int parity,
                                 const SorState *st) {
  hipLaunchKernelGGL((k_sor_pair_rollback<double, TI, TJ, NT, MINW, CREG>), g, blk, 0, s, a, gx, gy, parity, st);
}

template <int TI, int TJ, int NT, int MINW, bool CREG>
static void launch_pair_batch_rollback(dim3 g, dim3 blk, hipStream_t s, const BatchBlock *tab, int nblk) {
  hipLaunchKernelGGL((k_sor_pair_batch_rollback<double, TI, TJ, NT, MINW, CREG>), g, blk, 0, s, tab, nblk);
}

// FV(tile_i, tile_j, threads, prefetch depth, waves per SIMD the registers
// must allow, iterations per pass): SorTile; PV(tile_i, tile_j, threads,
// waves per SIMD, iterations per pass, coefficients in registers):
// SorPairTile (prefetch depth 1; family 1 with register coefficients, 2 with
// LDS coefficients)
#define FV(TI, TJ, NT, D, MINB, NIT) \
  { TI, TJ, NT, D, MINB, NIT, launch_fused<TI, TJ, NT, D, MINB, NIT>, launch_fused<TI, TJ, NT, D, MINB, 1>, 0, \
    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr }
#define PV(TI, TJ, NT, MINW, NIT, CREG) \
  { TI, TJ, NT, 1, MINW, NIT, launch_pair<TI, TJ, NT, MINW, NIT, CREG>, launch_pair<TI, TJ, NT, MINW, 1, CREG>, \
    CREG ? 1 : 2, launch_pair_batch<TI, TJ, NT, MINW, NIT, CREG>, launch_pair_batch<TI, TJ, NT, MINW, 1, CREG>, \
    NIT == 2 ? launch_pair<TI, TJ, NT, MINW, NIT, CREG, true> : nullptr, \
    NIT == 2 ? launch_pair_batch<TI, TJ, NT, MINW, NIT, CREG, true> : nullptr, \
    NIT == 2 ? launch_pair_rollback<TI, TJ, NT, MINW, CREG> : nullptr, \
    NIT == 2 ? launch_pair_batch_rollback<TI, TJ, NT, MINW, CREG> : nullptr }
static const FusedVariant kFused[] = {
    {0, 0, 0, 0, 0, 1, nullptr, nullptr, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr},
    FV(64, 16, 256, 1, 1, 1),
    FV(64, 8, 256, 1, 1, 1),
    FV(64, 8, 256, 2, 1, 1),
    FV(128, 16, 512, 1, 1, 1),
    FV(128, 16, 512, 2, 1, 1),
    FV(64, 32, 512, 1, 1, 1),
    FV(64, 32, 512, 2, 1, 1),
    FV(128, 8, 256, 2, 1, 1),
    FV(64, 8, 256, 1, 4, 1),
    FV(64, 16, 256, 1, 1, 2),
    FV(64, 16, 512, 1, 1, 2),
    FV(64, 32, 512, 1, 1, 2),
    FV(64, 8, 256, 1, 1, 2),
    FV(32, 16, 256, 1, 1, 2),
    FV(64, 8, 256, 1, 2, 2),
    FV(64, 32, 1024, 1, 1, 1),
    FV(64, 32, 1024, 1, 1, 2),
    FV(128, 16, 1024, 1, 1, 1),
    FV(64, 16, 1024, 1, 1, 1),
    FV(128, 32, 1024, 1, 1, 1),
    PV(64, 32, 512, 1, 2, 1),
    PV(64, 32, 512, 1, 1, 1),
    PV(64, 32, 512, 4, 1, 0),
    PV(64, 28, 512, 4, 1, 0),
    PV(64, 24, 512, 4, 1, 0),
    PV(64, 32, 512, 1, 2, 0),
    PV(64, 32, 1024, 1, 2, 0),
    PV(64, 32, 1024, 1, 2, 1),
    PV(128, 16, 1024, 1, 2, 0),
    PV(64, 32, 1024, 1, 1, 0),
    PV(64, 22, 512, 1, 2, 1),
    PV(64, 22, 512, 1, 2, 0),
    PV(32, 32, 512, 1, 2, 1),
    PV(32, 32, 512, 1, 2, 0),
};
#undef FV
#undef PV
constexpr int kNumSorVariants = (int)(sizeof(kFused) / sizeof(kFused[0]));
static int g_sor_variant = 26;  // the fastest at 256^3 on MI355X (pair-owned 64x32, LDS coefficients, 2 iterations per pass)
static int g_sor_planes = 32;

// Stopping test of the FP64 SOR calls.  rule 0 (default): the reference's,
// max |pc - pco| < epsi tested every fifth iteration (PcorVcor.F90:351-375).
// rule 1: the FP64 residual, max |b/dt - A pc| over the fluid cells < epsi,
// tested every `cadence` iterations (and after the last of the budget);
// nIterPcor and the pcItaMax exit are the reference's in both.
struct SorStop {
  int rule = 0;
  int cadence = 5;
};
static SorStop g_stop;
static int check_period() { return g_stop.rule == 1 ? g_stop.cadence : 5; }

int64_t residual_check_launch(Block &b, const double *pc, const SorState *st);
int64_t residual_workgroups(const Block &b);

static int enqueue_split_iterations(Block &b, SorState *st, long long n, int64_t pc_ita_max, double epsi) {
  hipStream_t s = rt().stream;
  const int NXh = (b.d.NX + 1) / 2;
  dim3 blk(64, 4, 1);
  dim3 grd(div_up((b.d.nx + 1) / 2, blk.x), div_up(b.d.ny, blk.y), div_up(b.d.nz, kSorKz));
  const int64_t nwg = (int64_t)grd.x * grd.y * grd.z;
  for (long long q = 0; q < n; ++q) {
    hipLaunchKernelGGL(k_sor_split, grd, blk, 0, s, b.d, NXh, 1, b.bs[1], b.pcs[1], b.pcs[0], b.A[0], b.A[1],
                       b.A[2], b.omega, st, b.partials);
    hipLaunchKernelGGL(k_sor_split, grd, blk, 0, s, b.d, NXh, 0, b.bs[0], b.pcs[0], b.pcs[1], b.A[0], b.A[1],
                       b.A[2], b.omega, st, b.partials + nwg);
    hipLaunchKernelGGL(k_sor_iter_end, dim3(1), dim3(1024), 0, s, st, (long long)pc_ita_max, epsi, b.partials,
                       2 * nwg);
  }
  BCFD_CHECK_LAUNCH();
  return 0;
}

static void fused_grid(const Block &b, const FusedVariant &v, int planes, int &gx, int &gy, int &gz) {
  gx = (int)div_up(b.d.nx, v.ti);
  gy = (int)div_up(b.d.ny, v.tj);
  gz = (int)div_up(b.d.nz, planes);
}

// Passes of a call: pass q reads pc (q even) or pc2 and writes the other.
// A pass runs up to `nit` iterations but always ends at a check iteration
// (every fifth, where the reference tests the update norm) and at the end of
// the call's budget of `total` iterations, so the stopping test sees exactly
// the reference's iterates.  Enqueues passes until `chunk` more iterations
// are queued or the budget is spent; `done_iters` and `q` carry over.
// The reference's rule (update norm every fifth iteration) with two-iteration
// pair-owned passes: no pass is cut short for a check iteration.  A check
// iteration that is a pass's first has its norm taken in that pass; if it
// stops the call, the pass's second iteration is undone by redoing the first
// alone from the pass's source (rollback_pass), which the pass did not
// overwrite.  The result is that of one-iteration passes, bit for bit.
static bool two_iteration_schedule(const FusedVariant &v) {
  return g_stop.rule == 0 && v.nit == 2 && v.family != 0 && v.launch_first;
}

static int rollback_pass(Block &b, SorState *st, int variant, int planes) {
  const FusedVariant &v = kFused[variant];
  if (!two_iteration_schedule(v)) return 0;
  hipStream_t s = rt().stream;
  int gx, gy, gz;
  fused_grid(b, v, planes, gx, gy, gz);
  FusedArgs a;
  a.d = b.d;
  a.NXh = (b.d.NX + 1) / 2;
  a.KS = planes;
  a.zero_src = 0;
  a.bred = b.bs[1];
  a.bblack = b.bs[0];
  a.Ax = b.A[0];
  a.Ay = b.A[1];
  a.Az = b.A[2];
  a.omega = b.omega;
  for (int par = 0; par < 2; ++par) {
    a.src = par ? b.pc2 : b.f[BF_PC];
    a.dst = par ? b.f[BF_PC] : b.pc2;
    v.rollback(dim3((unsigned)((int64_t)gx * gy * gz)), dim3(v.nt), s, a, gx, gy, par, st);
  }
  hipLaunchKernelGGL(k_sor_rollback_end, dim3(1), dim3(64), 0, s, st);
  BCFD_CHECK_LAUNCH();
  return 0;
}

// (end_it, end_q): the iteration and pass count of the call's last end
// kernel; with `rel` the end kernels take their counts relative to it
// (k_sor_fused_end_rel), as a captured chunk needs.
static int enqueue_fused_passes(Block &b, SorState *st, int variant, int planes, long long start, long long total,
                                long long chunk, long long &done_iters, long long &q, int64_t pc_ita_max,
                                double epsi, bool rel, long long &end_it, long long &end_q) {
  hipStream_t s = rt().stream;
  const FusedVariant &v = kFused[variant];
  int gx, gy, gz;
  fused_grid(b, v, planes, gx, gy, gz);
  const int64_t nwg = (int64_t)gx * gy * gz;
  FusedArgs a;
  a.d = b.d;
  a.NXh = (b.d.NX + 1) / 2;
  a.KS = planes;
  a.zero_src = 0;
  a.bred = b.bs[1];
  a.bblack = b.bs[0];
  a.Ax = b.A[0];
  a.Ay = b.A[1];
  a.Az = b.A[2];
  a.omega = b.omega;
  const long long stop = done_iters + chunk < total ? done_iters + chunk : total;
  const long long C = check_period();
  const bool first_checks = two_iteration_schedule(v);
  while (done_iters < stop) {
    const long long it = start + done_iters + 1;
    long long L = v.nit;
    if (L > total - done_iters) L = total - done_iters;
    int check;
    if (first_checks) {
      // every pass two iterations (the budget's end aside); a check iteration
      // may be a pass's first (check 2) or last (check 1)
      check = (it + L - 1) % C == 0 ? 1 : (L == 2 && it % C == 0 ? 2 : 0);
    } else {
      const long long to_check = (C - it % C) % C + 1;  // up to and including the next check iteration
      if (L > to_check) L = to_check;
      check = (it + L - 1) % C == 0;
    }
    const long long last_it = it + L - 1;
    a.src = (q & 1) ? b.pc2 : b.f[BF_PC];
    a.dst = (q & 1) ? b.f[BF_PC] : b.pc2;
    // rule 1 takes no update norm: the residual of the pass's result is
    // reduced after it instead
    if (check == 2)
      v.launch_first(dim3((unsigned)nwg), dim3(v.nt), s, a, gx, gy, 1, st, b.partials);
    else
      (L == v.nit ? v.launch : v.launch1)(dim3((unsigned)nwg), dim3(v.nt), s, a, gx, gy,
                                          g_stop.rule == 0 ? check : 0, st, b.partials);
    done_iters += L;
    ++q;
    if (g_stop.rule == 1 && (check || done_iters == total)) {
      const int64_t nres = residual_check_launch(b, a.dst, st);
      if (rel)
        hipLaunchKernelGGL(k_sor_fused_end_rel, dim3(1), dim3(1024), 0, s, st, last_it - end_it, q - end_q, 1,
                           (long long)pc_ita_max, epsi, b.partials, nres, 0);
      else
        hipLaunchKernelGGL(k_sor_fused_end, dim3(1), dim3(1024), 0, s, st, last_it, start, q, 1,
                           (long long)pc_ita_max, epsi, b.partials, nres, 0);
      end_it = last_it;
      end_q = q;
    } else if (check || done_iters == total) {
      if (rel)
        hipLaunchKernelGGL(k_sor_fused_end_rel, dim3(1), dim3(1024), 0, s, st, last_it - end_it, q - end_q,
                           check != 0, (long long)pc_ita_max, epsi, b.partials, nwg, check == 2 ? 1 : 0);
      else
        hipLaunchKernelGGL(k_sor_fused_end, dim3(1), dim3(1024), 0, s, st, last_it, start, q, check != 0,
                           (long long)pc_ita_max, epsi, b.partials, nwg, check == 2 ? 1 : 0);
      end_it = last_it;
      end_q = q;
    }
  }
  BCFD_CHECK_LAUNCH();
  return 0;
}

// b/dt of the fluid cells into the colour-split FP64 arrays (marker elsewhere)
int split_bdt(Block &b, double dt) {
  if (ensure_split(b)) return -1;
  const int NXh = (b.d.NX + 1) / 2;
  dim3 fb(64, 4, 1), fg(div_up(b.d.NX, 64), div_up(b.d.NY, 4), (unsigned)b.d.NZ);
  hipLaunchKernelGGL(k_split_bdt, fg, fb, 0, rt().stream, b.d, NXh, b.f[BF_B], b.tag[BT_CELL], dt, b.bs[1], b.bs[0]);
  BCFD_CHECK_LAUNCH();
  return 0;
}

// buffers and the per-call set-up of the SOR iterations: b/dt with the
// not-fluid marker, and pc copied to its partner array (fused) or split
static int sor_prepare(Block &b, int variant, int planes, double dt) {
  hipStream_t s = rt().stream;
  if (ensure_split(b)) return -1;
  const int NXh = (b.d.NX + 1) / 2;
  dim3 fb(64, 4, 1), fg(div_up(b.d.NX, 64), div_up(b.d.NY, 4), (unsigned)b.d.NZ);
  if (variant > 0) {
    if (b.pc2_n != b.n2) {
      if (dev_alloc(&b.pc2, b.n2)) return -1;
      b.pc2_n = b.n2;
    }
    int gx, gy, gz;
    fused_grid(b, kFused[variant], planes, gx, gy, gz);
    if (grow(&b.partials, &b.cap_partials, (int64_t)gx * gy * gz)) return -1;
    if (g_stop.rule == 1 && grow(&b.partials, &b.cap_partials, residual_workgroups(b))) return -1;
    BCFD_HIP(hipMemcpyAsync(b.pc2, b.f[BF_PC], sizeof(double) * (size_t)b.n2, hipMemcpyDeviceToDevice, s));
  } else {
    hipLaunchKernelGGL(k_split, fg, fb, 0, s, b.d, NXh, b.f[BF_PC], b.pcs[1], b.pcs[0]);
  }
  hipLaunchKernelGGL(k_split_bdt, fg, fb, 0, s, b.d, NXh, b.f[BF_B], b.tag[BT_CELL], dt, b.bs[1], b.bs[0]);
  BCFD_CHECK_LAUNCH();
  return 0;
}

// the iterate back into pc after the call's iterations
static int sor_finish(Block &b, const SorState *st, int variant) {
  hipStream_t s = rt().stream;
  const int NXh = (b.d.NX + 1) / 2;
  dim3 fb(64, 4, 1), fg(div_up(b.d.NX, 64), div_up(b.d.NY, 4), (unsigned)b.d.NZ);
  if (variant > 0)
    hipLaunchKernelGGL(k_sor_finish, dim3(1024), dim3(256), 0, s, st, b.f[BF_PC], b.pc2, b.n2);
  else
    hipLaunchKernelGGL(k_merge, fg, fb, 0, s, b.d, NXh, b.f[BF_PC], b.pcs[1], b.pcs[0]);
  BCFD_CHECK_LAUNCH();
  return 0;
}

// queue about `chunk` more iterations of the call (see enqueue_fused_passes)
static int enqueue_iterations(Block &b, SorState *st, int variant, int planes, long long start, long long total,
                              long long chunk, long long &done_iters, long long &q, int64_t pc_ita_max,
                              double epsi, long long &end_it, long long &end_q) {
  if (variant > 0)
    return enqueue_fused_passes(b, st, variant, planes, start, total, chunk, done_iters, q, pc_ita_max, epsi, false,
                                end_it, end_q);
  const long long n = chunk < total - done_iters ? chunk : total - done_iters;
  done_iters += n;
  return enqueue_split_iterations(b, st, n, pc_ita_max, epsi);
}

int sor_timing_on() { return g_tm.on ? 1 : 0; }

// HIP-event timing of one SOR call (when enabled): begin records e0
int sor_timing_begin(hipEvent_t *e0) {
  *e0 = nullptr;
  if (!g_tm.on) return 0;
  BCFD_HIP(hipEventCreate(e0));
  BCFD_HIP(hipEventRecord(*e0, rt().stream));
  return 0;
}

// end records e1 and the iterations the call ran (nIterPcor - start)
int sor_timing_end(hipEvent_t e0, const Block &b, long long start) {
  if (!g_tm.on) return 0;
  hipStream_t s = rt().stream;
  hipEvent_t e1 = nullptr;
  BCFD_HIP(hipEventCreate(&e1));
  BCFD_HIP(hipEventRecord(e1, s));
  SorState *h = (SorState *)rt().h_pinned;
  BCFD_HIP(hipMemcpyAsync(h, b.sor, sizeof(SorState), hipMemcpyDeviceToHost, s));
  BCFD_HIP(hipStreamSynchronize(s));
  g_tm.ev.emplace_back(e0, e1);
  g_tm.iters += h->n_iter - start;
  g_tm.cells += (h->n_iter - start) * (int64_t)b.d.nx * b.d.ny * b.d.nz;
  return 0;
}

// the same for one call on each block of a set that ran together from
// iteration `start`
int sor_timing_end_set(hipEvent_t e0, const std::vector<Block *> &blocks, const std::vector<long long> &start) {
  if (!g_tm.on) return 0;
  hipStream_t s = rt().stream;
  hipEvent_t e1 = nullptr;
  BCFD_HIP(hipEventCreate(&e1));
  BCFD_HIP(hipEventRecord(e1, s));
  g_tm.ev.emplace_back(e0, e1);
  SorState *h = (SorState *)rt().h_pinned;
  for (size_t i = 0; i < blocks.size(); ++i) {
    const Block *b = blocks[i];
    BCFD_HIP(hipMemcpyAsync(h, b->sor, sizeof(SorState), hipMemcpyDeviceToHost, s));
    BCFD_HIP(hipStreamSynchronize(s));
    g_tm.iters += h->n_iter - start[i];
    g_tm.cells += (h->n_iter - start[i]) * (int64_t)b->d.nx * b->d.ny * b->d.nz;
  }
  return 0;
}

// Host polls of a set of SOR calls without draining the queue (as sor_poll
// for one): the states after chunk k are copied to pinned set k & 1 and read
// once chunk k+1 is queued; stop = every block had stopped at chunk k - 1.
static SorState *g_set_poll_h = nullptr;   // two sets of kMaxBlocks + 1 states
static hipEvent_t g_set_poll_ev[2] = {nullptr, nullptr};
// the set's states are gathered on the device (one small kernel) and read
// back with one copy, instead of one copy per block
static SorState **g_poll_ptrs = nullptr;   // kMaxBlocks + 1 device pointers
static std::vector<SorState *> g_poll_ptrs_h;
static SorState *g_poll_stage = nullptr;   // two sets of kMaxBlocks + 1 states

__global__ void k_gather_states(SorState *const *__restrict__ ptrs, int n, SorState *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = *ptrs[i];
}

// the gather's buffers and the device copy of the set's state pointers
static int gather_ready(const std::vector<Block *> &blocks) {
  hipStream_t s = rt().stream;
  if (!g_set_poll_h)
    BCFD_HIP(hipHostMalloc((void **)&g_set_poll_h, sizeof(SorState) * 2 * (kMaxBlocks + 1), hipHostMallocDefault));
  if (!g_poll_stage && (dev_alloc(&g_poll_stage, 2 * (kMaxBlocks + 1)) || dev_alloc(&g_poll_ptrs, kMaxBlocks + 1)))
    return -1;
  std::vector<SorState *> ptrs(blocks.size());
  for (size_t i = 0; i < blocks.size(); ++i) ptrs[i] = blocks[i]->sor;
  if (ptrs != g_poll_ptrs_h) {
    // a gather queued earlier may still read the old pointers
    BCFD_HIP(hipStreamSynchronize(s));
    g_poll_ptrs_h = ptrs;
    BCFD_HIP(hipMemcpyAsync(g_poll_ptrs, g_poll_ptrs_h.data(), sizeof(SorState *) * ptrs.size(),
                            hipMemcpyHostToDevice, s));
    BCFD_HIP(hipStreamSynchronize(s));
  }
  return 0;
}

// the SOR states of a set of blocks now (one gather, one copy, one wait)
int read_states(const std::vector<Block *> &blocks, std::vector<SorState> &out) {
  hipStream_t s = rt().stream;
  if (gather_ready(blocks)) return -1;
  const int n = (int)blocks.size();
  hipLaunchKernelGGL(k_gather_states, dim3(div_up(n, 64)), dim3(64), 0, s, g_poll_ptrs, n, g_poll_stage);
  BCFD_CHECK_LAUNCH();
  BCFD_HIP(hipMemcpyAsync(g_set_poll_h, g_poll_stage, sizeof(SorState) * (size_t)n, hipMemcpyDeviceToHost, s));
  BCFD_HIP(hipStreamSynchronize(s));
  out.assign(g_set_poll_h, g_set_poll_h + n);
  return 0;
}

int set_poll(const std::vector<Block *> &blocks, long long &k, bool &stop) {
  hipStream_t s = rt().stream;
  if (gather_ready(blocks)) return -1;
  for (auto &e : g_set_poll_ev)
    if (!e) BCFD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  const int cur = (int)(k & 1);
  const int n = (int)blocks.size();
  SorState *set = g_set_poll_h + cur * (kMaxBlocks + 1), *stage = g_poll_stage + cur * (kMaxBlocks + 1);
  hipLaunchKernelGGL(k_gather_states, dim3(div_up(n, 64)), dim3(64), 0, s, g_poll_ptrs, n, stage);
  BCFD_CHECK_LAUNCH();
  BCFD_HIP(hipMemcpyAsync(set, stage, sizeof(SorState) * (size_t)n, hipMemcpyDeviceToHost, s));
  BCFD_HIP(hipEventRecord(g_set_poll_ev[cur], s));
  stop = false;
  if (k > 0) {
    BCFD_HIP(hipEventSynchronize(g_set_poll_ev[cur ^ 1]));
    const SorState *prev = g_set_poll_h + (cur ^ 1) * (kMaxBlocks + 1);
    stop = true;
    for (size_t i = 0; i < blocks.size(); ++i) stop = stop && prev[i].done != 0;
  }
  ++k;
  return 0;
}

// the device's compute units (workgroups one launch needs to fill it)
static int device_cus() {
  static int cus = 0;
  if (!cus) {
    int v = 0;
    cus = hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, rt().device) == hipSuccess && v > 0 ? v
                                                                                                           : 256;
  }
  return cus;
}

// planes per workgroup for a set of blocks: the configured count, halved
// (down to 8) while their tiles would leave compute units without a
// workgroup; the choice changes the speed, never the result
static int planes_for(const std::vector<Block *> &blocks, int variant, int planes) {
  if (variant <= 0) return planes;
  const FusedVariant &v = kFused[variant];
  int p = planes;
  while (p > 8) {
    int64_t tiles = 0;
    for (const Block *b : blocks) {
      int gx, gy, gz;
      fused_grid(*b, v, p, gx, gy, gz);
      tiles += (int64_t)gx * gy * gz;
    }
    if (tiles >= device_cus()) break;
    p /= 2;
  }
  return p;
}

// Host polls of a SOR call's state without draining the queue: the state
// after chunk k is copied to a pinned slot and read only once chunk k+1 is
// queued, so the GPU has work while the host waits.  A chunk queued past the
// stop runs as launches that see the done flag and return at once.
static hipEvent_t g_poll_ev[2] = {nullptr, nullptr};

int sor_poll(const SorState *st, long long &k, bool &stop) {
  hipStream_t s = rt().stream;
  SorState *slots = (SorState *)rt().h_pinned + 1;  // slots 1 and 2 (slot 0: the synchronous reads)
  for (auto &e : g_poll_ev)
    if (!e) BCFD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  const int cur = (int)(k & 1);
  BCFD_HIP(hipMemcpyAsync(&slots[cur], st, sizeof(SorState), hipMemcpyDeviceToHost, s));
  BCFD_HIP(hipEventRecord(g_poll_ev[cur], s));
  stop = false;
  if (k > 0) {
    BCFD_HIP(hipEventSynchronize(g_poll_ev[cur ^ 1]));
    stop = slots[cur ^ 1].done != 0;
  }
  ++k;
  return 0;
}

// HIP graphs of SOR passes.  Between two host polls a call's passes follow
// a pattern fixed by the iteration number modulo the check period (where
// the check iterations fall, which passes are cut short), the pass parity
// and the iterations and passes since the last end kernel; with end
// kernels that take their counts relative to the previous one, the
// launches of such a chunk do not depend on where it sits in the call.
// The first chunk of each pattern is captured (the same launches, in the
// same order, as the eager path would issue) and instantiated, and later
// chunks of that pattern replay it: one launch instead of some 20 to 260.
// A chunk never contains the call's last iteration (whose end kernel the
// eager path issues), and a graph is kept only while the block's buffers
// it names stay where they were.
static int g_sor_graphs = 1;
static long long g_graph_captures = 0, g_graph_replays = 0;

// iterations per graph chunk: whole check periods, at least 40 iterations
static long long graph_span() {
  const long long C = check_period();
  const long long P = g_stop.rule == 0 ? 10 : C;   // two-iteration passes against every fifth iteration
  return P * ((40 + P - 1) / P);
}

static bool same_ptrs(const SorGraph &g, const Block &b) {
  const void *now[SorGraph::kPtrs] = {b.f[BF_PC], b.pc2, b.partials, b.bs[0], b.bs[1], b.A[0], b.A[1], b.A[2], b.sor};
  for (int q = 0; q < SorGraph::kPtrs; ++q)
    if (now[q] != g.ptrs[q]) return false;
  return true;
}

int sor_graphs_on() { return g_sor_graphs; }

void sor_graphs_release(Block &b) {
  for (SorGraph &g : b.sor_graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  b.sor_graphs.clear();
}

// one chunk of a call through its graph (captured on first use): advances
// done_iters, q and the end-kernel counts as the eager path would
static int graph_chunk(Block &b, int variant, int planes, long long start, long long total, long long &done_iters,
                       long long &q, int64_t pc_ita_max, double epsi, long long &end_it, long long &end_q) {
  hipStream_t s = rt().stream;
  const long long G = graph_span(), P = g_stop.rule == 0 ? 10 : check_period();
  const long long it0 = start + done_iters;   // iterations done before the chunk
  long long epsi_bits;
  memcpy(&epsi_bits, &epsi, sizeof epsi_bits);
  const long long key[SorGraph::kKey] = {variant, planes, g_stop.rule, check_period(), it0 % P, q & 1,
                                         it0 - end_it, q - end_q, (long long)pc_ita_max, epsi_bits};
  SorGraph *hit = nullptr;
  for (SorGraph &g : b.sor_graphs)
    if (!memcmp(g.key, key, sizeof key)) { hit = &g; break; }
  if (hit && !same_ptrs(*hit, b)) {
    sor_graphs_release(b);
    hit = nullptr;
  }
  if (!hit) {
    long long d2 = done_iters, q2 = q, e_it = end_it, e_q = end_q;
    BCFD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_fused_passes(b, b.sor, variant, planes, start, total, G, d2, q2, pc_ita_max, epsi, true,
                                        e_it, e_q);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return -1;
    }
    BCFD_HIP(ec);
    SorGraph g;
    memcpy(g.key, key, sizeof key);
    const void *now[SorGraph::kPtrs] = {b.f[BF_PC], b.pc2, b.partials, b.bs[0], b.bs[1], b.A[0], b.A[1], b.A[2], b.sor};
    memcpy(g.ptrs, now, sizeof now);
    g.d_iters = d2 - done_iters;
    g.d_q = q2 - q;
    g.end_it = e_it - it0;
    g.end_q = e_q - q;
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    BCFD_HIP(ei);
    b.sor_graphs.push_back(g);
    hit = &b.sor_graphs.back();
    ++g_graph_captures;
  }
  BCFD_HIP(hipGraphLaunch(hit->exec, s));
  ++g_graph_replays;
  end_it = it0 + hit->end_it;
  end_q = q + hit->end_q;
  done_iters += hit->d_iters;
  q += hit->d_q;
  return 0;
}

int redblack_sor(Block &b, int64_t pc_ita_max, double dt, double epsi) {
  if (precision_mode() == 1) return mixed_sor(b, pc_ita_max, dt);
  hipStream_t s = rt().stream;
  hipEvent_t e0 = nullptr;
  const int variant = b.sor_variant >= 0 ? b.sor_variant : g_sor_variant;
  const int planes = b.sor_variant >= 0 ? b.sor_planes : planes_for({&b}, variant, g_sor_planes);
  if (g_stop.rule == 1 && variant == 0) {
    set_error("redblack_sor: the residual stopping rule needs a fused SOR variant (>= 1), not variant 0");
    return -1;
  }
  if (sor_prepare(b, variant, planes, dt)) return -1;
  hipLaunchKernelGGL(k_sor_reset_call, dim3(1), dim3(1), 0, s, b.sor, 0);
  BCFD_CHECK_LAUNCH();
  SorState *h = (SorState *)rt().h_pinned;
  BCFD_HIP(hipMemcpyAsync(h, b.sor, sizeof(SorState), hipMemcpyDeviceToHost, s));
  BCFD_HIP(hipStreamSynchronize(s));
  long long start = h->n_iter;
  long long remaining = pc_ita_max + 1 - start;
  if (sor_timing_begin(&e0)) return -1;
  if (remaining < 1) remaining = 1;
  const long long total = remaining;
  const long long chunk = 50;
  long long done_iters = 0, q = 0, polls = 0, end_it = start, end_q = 0;
  while (done_iters < total) {
    // graph chunks once the call has issued an end kernel (the chunk's first
    // relative counts need one), ending before the call's last iteration
    const bool graph = g_sor_graphs && variant > 0 && done_iters > 0 && end_q > 0 &&
                       done_iters + graph_span() + 2 < total;
    if (graph) {
      if (graph_chunk(b, variant, planes, start, total, done_iters, q, pc_ita_max, epsi, end_it, end_q)) return -1;
    } else if (enqueue_iterations(b, b.sor, variant, planes, start, total, chunk, done_iters, q, pc_ita_max, epsi,
                                  end_it, end_q)) {
      return -1;
    }
    if (done_iters >= total) break;
    bool stop = false;
    if (sor_poll(b.sor, polls, stop)) return -1;
    if (stop) break;
  }
  if (variant > 0 && rollback_pass(b, b.sor, variant, planes)) return -1;
  if (sor_timing_end(e0, b, start)) return -1;
  return sor_finish(b, b.sor, variant);
}

// Batching: the SOR calls of several blocks (the fine blocks of
// poissonSolver, PcorVcor.F90:86-129, which the reference runs one block per
// OpenMP thread) in one launch per pass over a table of the blocks, for the
// pair-owned kernels; blocks whose calls start at the same iteration share a
// pass schedule.  Off: one block after the other.
static int g_sor_batch = 1;
// captured chunks of batched passes (block set in `blocks`)
struct BatchGraph {
  static constexpr int kKey = 12;
  long long key[kKey];
  std::vector<Block *> blocks;
  long long d_iters, d_q, end_done, end_q;
  hipGraphExec_t exec = nullptr;
};
static std::vector<BatchGraph> g_batch_graphs;

void batch_graphs_release() {
  for (BatchGraph &g : g_batch_graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  g_batch_graphs.clear();
}
static BatchBlock *g_batch_tab = nullptr;
static int64_t g_batch_cap = 0;
static double *g_batch_partials = nullptr;
static int64_t g_batch_pcap = 0;

// one SOR call on each of `blocks`, block i from iteration starts[i]; the
// starts agree modulo 5, so every block's check iterations (nIterPcor a
// multiple of 5, PcorVcor.F90:351) fall on the same passes.  A pass never
// spans the end of a block's budget: the set's schedule also breaks where a
// block's budget ends, and the end-of-pass kernel stops that block there.
static int batched_calls(const std::vector<Block *> &blocks, const std::vector<long long> &starts, int64_t pc_ita_max,
                         double epsi) {
  hipStream_t s = rt().stream;
  const int variant = g_sor_variant, planes = planes_for(blocks, variant, g_sor_planes);
  const FusedVariant &v = kFused[variant];
  const int nb = (int)blocks.size();
  std::vector<BatchBlock> tab((size_t)nb);
  std::vector<long long> budget((size_t)nb);
  long long total = 1;
  int64_t total_tiles = 0;
  for (int i = 0; i < nb; ++i) {
    Block &b = *blocks[i];
    int gx, gy, gz;
    fused_grid(b, v, planes, gx, gy, gz);
    BatchBlock &e = tab[(size_t)i];
    for (int par = 0; par < 2; ++par) {
      FusedArgs &a = e.a[par];
      a.d = b.d;
      a.NXh = (b.d.NX + 1) / 2;
      a.KS = planes;
      a.zero_src = 0;
      a.bred = b.bs[1];
      a.bblack = b.bs[0];
      a.Ax = b.A[0];
      a.Ay = b.A[1];
      a.Az = b.A[2];
      a.omega = b.omega;
      a.src = par ? b.pc2 : b.f[BF_PC];
      a.dst = par ? b.f[BF_PC] : b.pc2;
    }
    e.gx = gx;
    e.gy = gy;
    e.ntiles = gx * gy * gz;
    e.first = (int)total_tiles;
    e.st = b.sor;
    e.start = starts[(size_t)i];
    total_tiles += e.ntiles;
    // at least one iteration per call, as the reference's loop (PcorVcor.F90:309)
    budget[(size_t)i] = pc_ita_max + 1 - starts[(size_t)i] < 1 ? 1 : pc_ita_max + 1 - starts[(size_t)i];
    total = budget[(size_t)i] > total ? budget[(size_t)i] : total;
  }
  if (grow(&g_batch_tab, &g_batch_cap, nb) || grow(&g_batch_partials, &g_batch_pcap, total_tiles)) return -1;
  BCFD_HIP(hipMemcpyAsync(g_batch_tab, tab.data(), sizeof(BatchBlock) * nb, hipMemcpyHostToDevice, s));
  BCFD_HIP(hipStreamSynchronize(s));
  hipEvent_t e0 = nullptr;
  if (sor_timing_begin(&e0)) return -1;
  const long long s0 = starts[0];
  long long done_iters = 0, q = 0, polls = 0, end_done = 0, end_q = 0;
  const bool first_checks = two_iteration_schedule(v);
  long long min_budget = total;
  for (long long bt : budget) min_budget = bt < min_budget ? bt : min_budget;
  // passes until `upto` iterations of the set are queued; `rel`: end kernels
  // relative to the previous one (a captured chunk)
  auto passes = [&](long long upto, bool rel) -> int {
    while (done_iters < upto) {
      const long long it = s0 + done_iters + 1;
      long long L = v.nit;
      if (L > total - done_iters) L = total - done_iters;
      for (long long bt : budget)
        if (bt > done_iters && L > bt - done_iters) L = bt - done_iters;
      int check;
      if (first_checks) {
        check = (it + L - 1) % 5 == 0 ? 1 : (L == 2 && it % 5 == 0 ? 2 : 0);
      } else {
        const long long to_check = (5 - it % 5) % 5 + 1;  // up to and including the next check iteration
        if (L > to_check) L = to_check;
        check = (it + L - 1) % 5 == 0;
      }
      if (check == 2)
        v.batch_first(dim3((unsigned)total_tiles), dim3(v.nt), s, g_batch_tab, nb, (int)(q & 1), 1,
                      g_batch_partials);
      else
        (L == v.nit ? v.batch : v.batch1)(dim3((unsigned)total_tiles), dim3(v.nt), s, g_batch_tab, nb,
                                          (int)(q & 1), check, g_batch_partials);
      done_iters += L;
      ++q;
      bool ends = done_iters == total;
      for (long long bt : budget) ends = ends || bt == done_iters;
      if (check || ends) {
        if (rel)
          hipLaunchKernelGGL(k_sor_batch_end_rel, dim3((unsigned)nb), dim3(1024), 0, s, g_batch_tab,
                             done_iters - end_done, q - end_q, check != 0, (long long)pc_ita_max, epsi,
                             g_batch_partials, check == 2 ? 1 : 0);
        else
          hipLaunchKernelGGL(k_sor_batch_end, dim3((unsigned)nb), dim3(1024), 0, s, g_batch_tab, done_iters, q,
                             check != 0, (long long)pc_ita_max, epsi, g_batch_partials, check == 2 ? 1 : 0);
        end_done = done_iters;
        end_q = q;
      }
      BCFD_CHECK_LAUNCH();
    }
    return 0;
  };
  while (done_iters < total) {
    // graph chunks (as graph_chunk) once an end kernel has run, ending before
    // the earliest budget end of the set
    const long long G = graph_span();
    if (g_sor_graphs && end_q > 0 && done_iters + G + 2 < min_budget) {
      long long epsi_bits;
      memcpy(&epsi_bits, &epsi, sizeof epsi_bits);
      const long long key[BatchGraph::kKey] = {variant, planes, nb, total_tiles, (s0 + done_iters) % 10, q & 1,
                                               done_iters - end_done, q - end_q, (long long)pc_ita_max, epsi_bits,
                                               (long long)(intptr_t)g_batch_tab,
                                               (long long)(intptr_t)g_batch_partials};
      BatchGraph *hit = nullptr;
      for (BatchGraph &g : g_batch_graphs)
        if (!memcmp(g.key, key, sizeof key) && g.blocks == blocks) { hit = &g; break; }
      const long long d0 = done_iters, q0 = q;
      if (!hit) {
        BCFD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = passes(done_iters + G, true);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(s, &graph);
        if (rc) {
          if (graph) (void)hipGraphDestroy(graph);
          return -1;
        }
        BCFD_HIP(ec);
        BatchGraph g;
        memcpy(g.key, key, sizeof key);
        g.blocks = blocks;
        g.d_iters = done_iters - d0;
        g.d_q = q - q0;
        g.end_done = end_done - d0;
        g.end_q = end_q - q0;
        const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        BCFD_HIP(ei);
        g_batch_graphs.push_back(g);
        hit = &g_batch_graphs.back();
        ++g_graph_captures;
      }
      BCFD_HIP(hipGraphLaunch(hit->exec, s));
      ++g_graph_replays;
      done_iters = d0 + hit->d_iters;
      q = q0 + hit->d_q;
      end_done = d0 + hit->end_done;
      end_q = q0 + hit->end_q;
    } else if (passes(done_iters + 50 < total ? done_iters + 50 : total, false)) {
      return -1;
    }
    if (done_iters < total) {
      // every block stopped: no more passes (PcorVcor.F90:365-375)
      bool stop = false;
      if (set_poll(blocks, polls, stop)) return -1;
      if (stop) break;
    }
  }
  if (first_checks) {
    // blocks that stopped on a pass's first iteration redo it alone
    v.batch_rollback(dim3((unsigned)total_tiles), dim3(v.nt), s, g_batch_tab, nb);
    hipLaunchKernelGGL(k_sor_rollback_end_set, dim3(div_up(nb, 64)), dim3(64), 0, s, g_batch_tab, nb);
    BCFD_CHECK_LAUNCH();
  }
  if (sor_timing_end_set(e0, blocks, starts)) return -1;
  for (Block *b : blocks)
    if (sor_finish(*b, b->sor, variant)) return -1;
  return 0;
}

// the SOR calls of `blocks` (REDBLACKSOR_linear on each), batched where the
// kernel and the calls allow
int redblack_sor_set(const std::vector<Block *> &blocks, int64_t pc_ita_max, double dt, double epsi) {
  if (precision_mode() == 1 && g_sor_batch && mixed_batchable() && blocks.size() > 1)
    return mixed_sor_set(blocks, pc_ita_max, dt);
  const int variant = g_sor_variant;
  const bool can = g_sor_batch && precision_mode() == 0 && g_stop.rule == 0 && kFused[variant].batch &&
                   blocks.size() > 1;
  if (!can) {
    for (Block *b : blocks)
      if (redblack_sor(*b, pc_ita_max, dt, epsi)) return -1;
    return 0;
  }
  hipStream_t s = rt().stream;
  // set-up of every call, then the start iteration of each (nIterPcor so far)
  std::vector<long long> start(blocks.size());
  for (size_t i = 0; i < blocks.size(); ++i) {
    Block &b = *blocks[i];
    if (sor_prepare(b, variant, planes_for(blocks, variant, g_sor_planes), dt)) return -1;
    hipLaunchKernelGGL(k_sor_reset_call, dim3(1), dim3(1), 0, s, b.sor, 0);
    BCFD_CHECK_LAUNCH();
  }
  std::vector<SorState> st0;
  if (read_states(blocks, st0)) return -1;
  for (size_t i = 0; i < blocks.size(); ++i) start[i] = st0[i].n_iter;
  // blocks whose calls start at iterations equal modulo 5 share a pass
  // schedule (their check iterations coincide)
  std::vector<bool> taken(blocks.size(), false);
  for (size_t i = 0; i < blocks.size(); ++i) {
    if (taken[i]) continue;
    std::vector<Block *> group;
    std::vector<long long> gstart;
    for (size_t j = i; j < blocks.size(); ++j)
      if (!taken[j] && start[j] % 5 == start[i] % 5) {
        group.push_back(blocks[j]);
        gstart.push_back(start[j]);
        taken[j] = true;
      }
    if (batched_calls(group, gstart, pc_ita_max, epsi)) return -1;
  }
  return 0;
}

// `iters` SOR iterations of the given variant on the block's current pc and
// b, timed with HIP events; pc and the block's SOR state are left unchanged
static SorState *g_probe_state = nullptr;

int sor_time(Block &b, int variant, int planes, int iters, double dt, double *ms) {
  hipStream_t s = rt().stream;
  if (!g_probe_state) {
    if (dev_alloc(&g_probe_state, 1)) return -1;
  }
  double *backup = nullptr;
  if (dev_alloc(&backup, b.n2)) return -1;
  BCFD_HIP(hipMemcpyAsync(backup, b.f[BF_PC], sizeof(double) * (size_t)b.n2, hipMemcpyDeviceToDevice, s));
  int rc = sor_prepare(b, variant, planes, dt);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc) {
    hipLaunchKernelGGL(k_sor_reset_call, dim3(1), dim3(1), 0, s, g_probe_state, 1);
    rc = hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
         hipEventRecord(e0, s) != hipSuccess;
  }
  // pcItaMax out of reach and a negative tolerance: exactly `iters` iterations
  long long done_iters = 0, q = 0, end_it = 0, end_q = 0;
  if (!rc)
    rc = enqueue_iterations(b, g_probe_state, variant, planes, 0, iters, iters, done_iters, q, (int64_t)1 << 62,
                            -1.0, end_it, end_q);
  float t = 0.f;
  if (!rc) rc = hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t, e0, e1) != hipSuccess;
  if (rc) set_error("sor_time: variant %d planes %d failed", variant, planes);
  (void)hipMemcpyAsync(b.f[BF_PC], backup, sizeof(double) * (size_t)b.n2, hipMemcpyDeviceToDevice, s);
  (void)hipStreamSynchronize(s);
  (void)hipFree(backup);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  *ms = t;
  return rc ? -1 : 0;
}

int correct(Block &b, double dt) {
  hipStream_t s = rt().stream;
  dim3 blk(64, 4, 1);
  hipLaunchKernelGGL(k_correct, interior_grid(b.d, blk), blk, 0, s, b.d, b.tag[BT_CELL], b.f[BF_PC],
                     b.f[BF_P], b.f[BF_UT], b.f[BF_VT], b.f[BF_WT], b.g[BG_DX], b.g[BG_DY],
                     b.g[BG_DZ], dt);
  BCFD_CHECK_LAUNCH();
  return 0;
}

int velocity_bc(Block &b, double dt, double uc, double inlet) {
  hipStream_t s = rt().stream;
  hipLaunchKernelGGL(k_bc_x, dim3(div_up(b.d.ny, 64), b.d.nz), dim3(64), 0, s, b.d, b.f[BF_U],
                     b.f[BF_V], b.f[BF_W], b.f[BF_UT], b.f[BF_VT], b.f[BF_WT], b.g[BG_DX], dt, uc, inlet);
  hipLaunchKernelGGL(k_bc_y, dim3(div_up(b.d.nx, 64), b.d.nz), dim3(64), 0, s, b.d, b.f[BF_UT],
                     b.f[BF_VT], b.f[BF_WT]);
  hipLaunchKernelGGL(k_bc_z, dim3(div_up(b.d.nx, 64), b.d.ny), dim3(64), 0, s, b.d, b.f[BF_UT],
                     b.f[BF_VT], b.f[BF_WT]);
  BCFD_CHECK_LAUNCH();
  return 0;
}

int norm_copy(Block &b, double dt) {
  hipStream_t s = rt().stream;
  dim3 blk(64, 4, 1);
  dim3 grd(div_up(b.d.NX, blk.x), div_up(b.d.NY, blk.y), (unsigned)b.d.NZ);
  if (ensure_split(b)) return -1;
  hipLaunchKernelGGL(k_norm_copy, grd, blk, 0, s, b.d, b.tag[BT_CELL], b.f[BF_U], b.f[BF_V],
                     b.f[BF_W], b.f[BF_UT], b.f[BF_VT], b.f[BF_WT], dt, b.partials);
  hipLaunchKernelGGL(k_norm_end, dim3(1), dim3(1024), 0, s, b.sor, b.partials, (int64_t)grd.x * grd.y * grd.z);
  BCFD_CHECK_LAUNCH();
  return 0;
}

// b = pc = 0, computeDiv, n_iter SOR iterations of the FP64 method,
// correctPressure + correctVelocity (PcorVcor.F90:67-77, 222-377) on one
// block, no interface updates or boundary conditions
int projection(Block &b, int64_t n_iter, double dt) {
  if (check_ready(b)) return -1;
  if (n_iter < 1) { set_error("projection: iteration count %lld < 1", (long long)n_iter); return -1; }
  if (precision_mode() != 0) { set_error("projection: the FP64 method only (set_precision 0)"); return -1; }
  hipStream_t s = rt().stream;
  hipLaunchKernelGGL(k_zero_pcor, dim3(div_up(b.n2, 256)), dim3(256), 0, s, b.f[BF_B], b.f[BF_PC], b.n2);
  hipLaunchKernelGGL(k_sor_reset_call, dim3(1), dim3(1), 0, s, b.sor, 1);
  BCFD_CHECK_LAUNCH();
  if (compute_div(b)) return -1;
  // a negative tolerance never stops the iterations: nIterPcor > n_iter - 1 ends them after n_iter
  if (redblack_sor(b, n_iter - 1, dt, -1.0)) return -1;
  return correct(b, dt);
}

}  // namespace bcfd

using namespace bcfd;

extern "C" {

int biocfd_velocity_bc(int id, double dt, double uc) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (!b->g[BG_DX]) { set_error("grids not uploaded"); return -1; }
  return velocity_bc(*b, dt, uc, uc);
}

int biocfd_solid_cell_bc(int id) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  dim3 blk(64, 4, 1);
  hipLaunchKernelGGL(k_solid_bc, interior_grid(b->d, blk), blk, 0, rt().stream, b->d, b->tag[BT_CELL],
                     b->f[BF_UT], b->f[BF_VT], b->f[BF_WT], b->f[BF_P], b->f[BF_U], b->f[BF_V],
                     b->f[BF_W], 0);
  BCFD_CHECK_LAUNCH();
  return 0;
}

int biocfd_solid_cell_bc_move(int id, int64_t move_check) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (move_check != 1) return 0;
  dim3 blk(64, 4, 1);
  hipLaunchKernelGGL(k_solid_bc, interior_grid(b->d, blk), blk, 0, rt().stream, b->d, b->tag[BT_CELL],
                     b->f[BF_UT], b->f[BF_VT], b->f[BF_WT], b->f[BF_P], b->f[BF_U], b->f[BF_V],
                     b->f[BF_W], 1);
  BCFD_CHECK_LAUNCH();
  return 0;
}

int biocfd_update_velocity_newv(int id, int64_t move_check) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (move_check != 1) return 0;
  hipLaunchKernelGGL(k_copy3, dim3(div_up(b->n2, 256)), dim3(256), 0, rt().stream, b->f[BF_UT],
                     b->f[BF_VT], b->f[BF_WT], b->f[BF_U], b->f[BF_V], b->f[BF_W], b->n2);
  BCFD_CHECK_LAUNCH();
  return 0;
}

int biocfd_timing_enable(int on) {
  BCFD_API_LOCK();
  if (ensure_init()) return -1;
  for (auto &pr : g_tm.ev) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
  g_tm = SorTiming();
  g_tm.on = on != 0;
  return 0;
}

int biocfd_timing_read(double *sor_ms, int64_t *sor_iters, int64_t *sor_cell_iters) {
  BCFD_API_LOCK();
  if (ensure_init()) return -1;
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  double ms = 0.0;
  for (auto &pr : g_tm.ev) {
    float t = 0.f;
    BCFD_HIP(hipEventElapsedTime(&t, pr.first, pr.second));
    ms += t;
  }
  *sor_ms = ms;
  *sor_iters = g_tm.iters;
  *sor_cell_iters = g_tm.cells;
  return 0;
}

int biocfd_sor_batching(int on) {
  BCFD_API_LOCK();
  g_sor_batch = on ? 1 : 0;
  return 0;
}

int biocfd_sor_graphs(int on) {
  BCFD_API_LOCK();
  g_sor_graphs = on ? 1 : 0;
  return 0;
}

int biocfd_sor_graph_stats(int64_t *captures, int64_t *launches) {
  BCFD_API_LOCK();
  *captures = g_graph_captures;
  *launches = g_graph_replays;
  return 0;
}

int biocfd_sor_stop_rule(int rule, int cadence) {
  BCFD_API_LOCK();
  if (rule == 0) {
    g_stop.rule = 0;
    g_stop.cadence = 5;
    return 0;
  }
  if (rule != 1 || cadence < 1) {
    set_error("sor_stop_rule: rule %d cadence %d (rule 0 = update norm every 5th iteration, rule 1 = FP64 "
              "residual every cadence >= 1 iterations)", rule, cadence);
    return -1;
  }
  g_stop.rule = 1;
  g_stop.cadence = cadence;
  return 0;
}

int biocfd_sor_block_config(int blk, int variant, int planes_per_workgroup) {
  BCFD_API_LOCK();
  Block *b = get_block(blk);
  if (!b) return -1;
  if (variant < -1 || variant >= kNumSorVariants) {
    set_error("sor_block_config: variant %d outside -1..%d", variant, kNumSorVariants - 1);
    return -1;
  }
  if (variant >= 0 && (planes_per_workgroup < 1 || planes_per_workgroup > 4096)) {
    set_error("sor_block_config: planes per workgroup %d outside 1..4096", planes_per_workgroup);
    return -1;
  }
  b->sor_variant = variant;
  b->sor_planes = planes_per_workgroup;
  return 0;
}

int biocfd_sor_config(int variant, int planes_per_workgroup) {
  BCFD_API_LOCK();
  if (variant < 0 || variant >= kNumSorVariants) {
    set_error("sor_config: variant %d outside 0..%d", variant, kNumSorVariants - 1);
    return -1;
  }
  if (planes_per_workgroup < 1 || planes_per_workgroup > 4096) {
    set_error("sor_config: planes per workgroup %d outside 1..4096", planes_per_workgroup);
    return -1;
  }
  g_sor_variant = variant;
  g_sor_planes = planes_per_workgroup;
  return 0;
}

int biocfd_sor_variant_info(int variant, int info[7]) {
  BCFD_API_LOCK();
  if (variant < 0 || variant >= kNumSorVariants) {
    set_error("sor_variant_info: variant %d outside 0..%d", variant, kNumSorVariants - 1);
    return -1;
  }
  const FusedVariant &v = kFused[variant];
  info[0] = v.ti;
  info[1] = v.tj;
  info[2] = v.nt;
  info[3] = v.depth;
  info[4] = v.minb;
  info[5] = v.nit;
  info[6] = v.family;
  return 0;
}

int biocfd_sor_time(int id, int variant, int planes_per_workgroup, int iters, double dt, double *ms) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b || check_ready(*b)) return -1;
  if (variant < 0 || variant >= kNumSorVariants || planes_per_workgroup < 1 || planes_per_workgroup > 4096 ||
      iters < 1) {
    set_error("sor_time: bad variant %d / planes %d / iterations %d", variant, planes_per_workgroup, iters);
    return -1;
  }
  return sor_time(*b, variant, planes_per_workgroup, iters, dt, ms);
}

int biocfd_projection(int id, int64_t n_iter, double dt) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  return projection(*b, n_iter, dt);
}

int biocfd_compute_div(int id) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b || check_ready(*b)) return -1;
  return compute_div(*b);
}

int biocfd_redblack_sor(int id, int64_t pc_ita_max, double dt, double epsi) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b || check_ready(*b)) return -1;
  return redblack_sor(*b, pc_ita_max, dt, epsi);
}

int biocfd_correct(int id, double dt) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b || check_ready(*b)) return -1;
  return correct(*b, dt);
}

int biocfd_poisson_solver(int first, int last, int64_t pc_ita_max, double epsi, double dt, double uc) {
  BCFD_API_LOCK();
  if (ensure_init()) return -1;
  if (first < 1 || last < first || last > kMaxBlocks) {
    set_error("poisson_solver: bad block range %d..%d", first, last);
    return -1;
  }
  for (int g = first; g <= last; ++g) {
    Block *b = get_block(g);
    if (!b || check_ready(*b)) return -1;
  }
  hipStream_t s = rt().stream;
  for (int g = first; g <= last; ++g) {
    Block &b = rt().blk[g];
    hipLaunchKernelGGL(k_zero_pcor, dim3(div_up(b.n2, 256)), dim3(256), 0, s, b.f[BF_B], b.f[BF_PC], b.n2);
    hipLaunchKernelGGL(k_sor_reset_call, dim3(1), dim3(1), 0, s, b.sor, 1);
  }
  BCFD_CHECK_LAUNCH();
  if (interface_fine_update(1) || interface_coarse_update(1)) return -1;
  std::vector<Block *> fine;
  for (int g = first; g <= last; ++g) {
    Block &b = rt().blk[g];
    if (compute_div(b)) return -1;
    if (g != 1) fine.push_back(&b);
  }
  if (redblack_sor_set(fine, pc_ita_max, dt, epsi)) return -1;
  if (interface_coarse_update(2)) return -1;
  if (first == 1 && redblack_sor(rt().blk[1], pc_ita_max, dt, epsi)) return -1;
  if (interface_fine_update(2)) return -1;
  if (redblack_sor_set(fine, pc_ita_max, dt, epsi)) return -1;
  if (interface_coarse_update(2)) return -1;
  for (int g = first; g <= last; ++g)
    if (correct(rt().blk[g], dt)) return -1;
  if (first == 1 && velocity_bc(rt().blk[1], dt, uc, uc)) return -1;
  for (int g = first; g <= last; ++g)
    if (norm_copy(rt().blk[g], dt)) return -1;
  if (interface_fine_update(0) || interface_coarse_update(0)) return -1;
  return health_after_solve(first, last);
}

}  // extern "C"
