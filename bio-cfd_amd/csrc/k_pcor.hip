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
#include "biocfd_rt.hpp"

namespace bcfd {

int interface_coarse_update(int which);
int interface_fine_update(int which);

__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_xor(v, o, 64);
    v = v > w ? v : w;
  }
  return v;
}

// bit pattern of a non-negative double orders like its value
__device__ inline void atomic_max_nonneg(unsigned long long *dst, double v) {
  atomicMax(dst, (unsigned long long)__double_as_longlong(v));
}

__global__ void k_zero_pcor(double *__restrict__ b, double *__restrict__ pc, int64_t n) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) { b[q] = 0.0; pc[q] = 0.0; }
}

__global__ void k_sor_reset_call(SorState *st, int reset_iter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (reset_iter) { st->n_iter = 0; st->derr2 = 0.0; st->derr_std_st = 0.0; }
    st->derr4 = 11111.0;
    st->max_bits = 0ull;
    st->done = 0;
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

// One colour of REDBLACKSOR_linear (PcorVcor.F90:313-330 / 332-349).
// colour 1 = red ((i+j+k) odd), 0 = black.  Each thread owns one cell of the
// colour in row (j,k).
__global__ void k_sor_colour(Dims d, int colour, const uint8_t *__restrict__ cell,
                             const double *__restrict__ b, double *__restrict__ pc,
                             const double *__restrict__ Ax, const double *__restrict__ Ay,
                             const double *__restrict__ Az, double inv_dt_unused, double dt,
                             double omega, SorState *st) {
  (void)inv_dt_unused;
  if (st->done) return;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  const int i0 = 2 + ((colour - (2 + j + k)) & 1);
  const int i = i0 + 2 * (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const bool check = ((st->n_iter + 1) % 5) == 0;
  double delta = 0.0;
  if (j <= d.ny + 1 && i <= d.nx + 1) {
    const int64_t c = ix3(d, i, j, k);
    if (cell[c] == 0) {
      const int64_t sy = d.NX, sz = (int64_t)d.NX * d.NY;
      const int nx = d.nx, ny = d.ny, nz = d.nz;
      const double old = pc[c];
      const double gs = (b[c] / dt - Ax[(i - 2)] * pc[c - 1] - Ax[(i - 2) + 2 * nx] * pc[c + 1] -
                         Ay[(j - 2)] * pc[c - sy] - Ay[(j - 2) + 2 * ny] * pc[c + sy] -
                         Az[(k - 2)] * pc[c - sz] - Az[(k - 2) + 2 * nz] * pc[c + sz]) /
                        (Ax[(i - 2) + nx] + Ay[(j - 2) + ny] + Az[(k - 2) + nz]);
      const double nw = (1.0 - omega) * old + omega * gs;
      pc[c] = nw;
      delta = fabs(nw - old);
    }
  }
  if (check) {
    delta = wave_max(delta);
    if ((threadIdx.x + threadIdx.y * blockDim.x) % 64 == 0 && delta > 0.0)
      atomic_max_nonneg(&st->max_bits, delta);
  }
}

// end-of-iteration bookkeeping of REDBLACKSOR_linear (PcorVcor.F90:309, 351-375)
__global__ void k_sor_iter_end(SorState *st, long long pc_ita_max, double epsi) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || st->done) return;
  const long long it = st->n_iter + 1;
  st->n_iter = it;
  if (it % 5 == 0) {
    st->derr4 = __longlong_as_double((long long)st->max_bits);
    st->max_bits = 0ull;
  }
  st->derr2 = st->derr4;
  if (st->derr4 < epsi || it > pc_ita_max) st->done = 1;
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
                            const double *__restrict__ wt, double dt, SorState *st) {
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
  m = wave_max(m);
  if ((threadIdx.x + threadIdx.y * blockDim.x) % 64 == 0 && m > 0.0) atomic_max_nonneg(&st->std_bits, m);
}

__global__ void k_norm_begin(SorState *st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) st->std_bits = 0ull;
}
__global__ void k_norm_end(SorState *st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) st->derr_std_st = __longlong_as_double((long long)st->std_bits);
}

// velocityBC (BoundaryConditions.f90:11-74): three face passes in order
__global__ void k_bc_x(Dims d, const double *__restrict__ u, const double *__restrict__ v,
                       const double *__restrict__ w, double *__restrict__ ut, double *__restrict__ vt,
                       double *__restrict__ wt, const double *__restrict__ dxv, double dt, double uc) {
  const int j = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int k = 2 + blockIdx.y;
  if (j > d.ny + 1) return;
  const int nx = d.nx;
  const int64_t c1 = ix3(d, 1, j, k), cn = ix3(d, nx, j, k), cn1 = ix3(d, nx + 1, j, k),
                cn2 = ix3(d, nx + 2, j, k);
  ut[c1] = uc;
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
int redblack_sor(Block &b, int64_t pc_ita_max, double dt, double epsi) {
  hipStream_t s = rt().stream;
  hipLaunchKernelGGL(k_sor_reset_call, dim3(1), dim3(1), 0, s, b.sor, 0);
  BCFD_CHECK_LAUNCH();
  dim3 blk(64, 4, 1);
  const int half = (b.d.nx + 1) / 2 + 1;
  dim3 grd(div_up(half, blk.x), div_up(b.d.ny, blk.y), (unsigned)b.d.nz);
  SorState *h = (SorState *)rt().h_pinned;
  BCFD_HIP(hipMemcpyAsync(h, b.sor, sizeof(SorState), hipMemcpyDeviceToHost, s));
  BCFD_HIP(hipStreamSynchronize(s));
  long long start = h->n_iter;
  long long remaining = pc_ita_max + 1 - start;
  if (remaining < 1) remaining = 1;
  const long long chunk = 50;
  while (remaining > 0) {
    long long n = remaining < chunk ? remaining : chunk;
    for (long long q = 0; q < n; ++q) {
      hipLaunchKernelGGL(k_sor_colour, grd, blk, 0, s, b.d, 1, b.tag[BT_CELL], b.f[BF_B], b.f[BF_PC],
                         b.A[0], b.A[1], b.A[2], 0.0, dt, b.omega, b.sor);
      hipLaunchKernelGGL(k_sor_colour, grd, blk, 0, s, b.d, 0, b.tag[BT_CELL], b.f[BF_B], b.f[BF_PC],
                         b.A[0], b.A[1], b.A[2], 0.0, dt, b.omega, b.sor);
      hipLaunchKernelGGL(k_sor_iter_end, dim3(1), dim3(64), 0, s, b.sor, (long long)pc_ita_max, epsi);
    }
    BCFD_CHECK_LAUNCH();
    remaining -= n;
    if (remaining <= 0) break;
    BCFD_HIP(hipMemcpyAsync(h, b.sor, sizeof(SorState), hipMemcpyDeviceToHost, s));
    BCFD_HIP(hipStreamSynchronize(s));
    if (h->done) break;
  }
  return 0;
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

int velocity_bc(Block &b, double dt, double uc) {
  hipStream_t s = rt().stream;
  hipLaunchKernelGGL(k_bc_x, dim3(div_up(b.d.ny, 64), b.d.nz), dim3(64), 0, s, b.d, b.f[BF_U],
                     b.f[BF_V], b.f[BF_W], b.f[BF_UT], b.f[BF_VT], b.f[BF_WT], b.g[BG_DX], dt, uc);
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
  hipLaunchKernelGGL(k_norm_begin, dim3(1), dim3(1), 0, s, b.sor);
  hipLaunchKernelGGL(k_norm_copy, grd, blk, 0, s, b.d, b.tag[BT_CELL], b.f[BF_U], b.f[BF_V],
                     b.f[BF_W], b.f[BF_UT], b.f[BF_VT], b.f[BF_WT], dt, b.sor);
  hipLaunchKernelGGL(k_norm_end, dim3(1), dim3(1), 0, s, b.sor);
  BCFD_CHECK_LAUNCH();
  return 0;
}

}  // namespace bcfd

using namespace bcfd;

extern "C" {

int biocfd_velocity_bc(int id, double dt, double uc) {
  Block *b = get_block(id);
  if (!b) return -1;
  if (!b->g[BG_DX]) { set_error("grids not uploaded"); return -1; }
  return velocity_bc(*b, dt, uc);
}

int biocfd_solid_cell_bc(int id) {
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
  Block *b = get_block(id);
  if (!b) return -1;
  if (move_check != 1) return 0;
  hipLaunchKernelGGL(k_copy3, dim3(div_up(b->n2, 256)), dim3(256), 0, rt().stream, b->f[BF_UT],
                     b->f[BF_VT], b->f[BF_WT], b->f[BF_U], b->f[BF_V], b->f[BF_W], b->n2);
  BCFD_CHECK_LAUNCH();
  return 0;
}

int biocfd_compute_div(int id) {
  Block *b = get_block(id);
  if (!b || check_ready(*b)) return -1;
  return compute_div(*b);
}

int biocfd_redblack_sor(int id, int64_t pc_ita_max, double dt, double epsi) {
  Block *b = get_block(id);
  if (!b || check_ready(*b)) return -1;
  return redblack_sor(*b, pc_ita_max, dt, epsi);
}

int biocfd_correct(int id, double dt) {
  Block *b = get_block(id);
  if (!b || check_ready(*b)) return -1;
  return correct(*b, dt);
}

int biocfd_poisson_solver(int first, int last, int64_t pc_ita_max, double epsi, double dt, double uc) {
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
  for (int g = first; g <= last; ++g) {
    Block &b = rt().blk[g];
    if (compute_div(b)) return -1;
    if (g != 1 && redblack_sor(b, pc_ita_max, dt, epsi)) return -1;
  }
  if (interface_coarse_update(2)) return -1;
  if (first == 1 && redblack_sor(rt().blk[1], pc_ita_max, dt, epsi)) return -1;
  if (interface_fine_update(2)) return -1;
  for (int g = first; g <= last; ++g)
    if (g != 1 && redblack_sor(rt().blk[g], pc_ita_max, dt, epsi)) return -1;
  if (interface_coarse_update(2)) return -1;
  for (int g = first; g <= last; ++g)
    if (correct(rt().blk[g], dt)) return -1;
  if (first == 1 && velocity_bc(rt().blk[1], dt, uc)) return -1;
  for (int g = first; g <= last; ++g)
    if (norm_copy(rt().blk[g], dt)) return -1;
  if (interface_fine_update(0) || interface_coarse_update(0)) return -1;
  return 0;
}

}  // extern "C"
