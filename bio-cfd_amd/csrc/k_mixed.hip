// Mixed-precision pressure solve: the FP32-smoothing variant of
// REDBLACKSOR_linear (PcorVcor.F90:288-377) as FP64 defect correction.
//
//   r  = b/dt - A pc          FP64 arithmetic, stored as FP32 (colour-split),
//                             with its FP64 max-norm
//   e  = S^k(0; r)            k FP32 red-black SOR sweeps from e = 0 (the
//                             fused tile kernel of sor_fused.hpp, Real = float)
//   pc = pc + e               FP64
//
// until ||r||inf < tol or the sweep budget of the call (pcItaMax + 1 -
// nIterPcor, the FP64 method's iteration budget) is spent.  k sweeps of a
// stationary iteration from a zero guess on the residual equation are k
// steps of the same iteration on pc, so the iterates follow the FP64 SOR's up
// to the FP32 rounding of each correction; the residual is always recomputed
// from the FP64 pc, so those roundings do not accumulate.
//
// Traffic per cycle of k sweeps at n cells: k sweeps 4+4 read (the first
// reads no iterate), 4 written; update + residual (one pass) pc 8, e 4, b/dt 8
// read, pc 8, r 4 written: (32 + 12k - 4) n bytes.  The FP32 iterate pair and residual
// (3 x 4 B per cell, 201 MB at 256^3) fit in MI355X's 256 MB Infinity Cache.
#include "biocfd_rt.hpp"
#include "reduce.hpp"
#include "sor_kernel.hpp"

namespace bcfd {

// tile of a residual workgroup: TI x TJ columns of 1024 threads (rows ty
// and ty + 1024 / TI), marching up to kResPlanes planes: 256 x 8 (wide in
// i: each side-halo cell of a row costs a whole cache line of pc and of e)
// or, for blocks at most 128 cells wide, 64 x 32
constexpr int kResNT = 1024, kResPlanes = 32;
template <int TI>
struct ResTile {
  static constexpr int R = kResNT / TI;                 // rows per half
  static constexpr int TJ = 2 * R;
  static constexpr int PW = TI + 2, SH = TJ + 2;        // plane footprint with a one-cell halo
  static constexpr int Halo = 2 * PW + 2 * TJ;          // halo cells, one per thread < Halo
  static_assert(Halo <= kResNT, "one halo cell per thread");
};

// With kUpdate: pc' = pc + e (every interior cell; e is zero off the fluid)
// is written to pc_out and the residual is taken of pc'; without, of pc.
// r = b/dt - A pc' at interior fluid cells in FP64, stored as FP32 into the
// colour-split arrays (marker at the other interior cells) when kStore, with
// the per-workgroup max |r| (without kStore only the maxima are written: the
// FP64 method's residual stopping test).
//
// The workgroup marches its tile in k with the values of plane k (tile and
// halo) in one of two LDS planes: step k stores plane k+1 (loaded in step
// k-1) into the other, issues the loads of plane k+2 and of the rhs of k+1,
// and computes plane k; i and j neighbours come from LDS, k neighbours from
// registers.  HBM traffic per cell: pc 8, e 4, b/dt 8 read, pc' 8, r 4
// written; the halo is re-read from L2.
template <int TI, bool kUpdate, bool kStore = true>
__global__ void __launch_bounds__(kResNT) k_mx_resupd(Dims d, int NXh, int planes, const double *__restrict__ pc,
                                                   const float *__restrict__ e, double *__restrict__ pc_out,
                                                   const double *__restrict__ bred, const double *__restrict__ bblack,
                                                   const double *__restrict__ Ax, const double *__restrict__ Ay,
                                                   const double *__restrict__ Az, float *__restrict__ rred,
                                                   float *__restrict__ rblack, const SorState *__restrict__ st,
                                                   double *__restrict__ partials) {
  if (st->done) return;
  using T = ResTile<TI>;
  constexpr int kResPW = T::PW, kResSH = T::SH, kResHalo = T::Halo, R = T::R;
  __shared__ double pl[2][kResSH][kResPW];
  const int tid = threadIdx.x, tx = tid % TI, ty = tid / TI;
  const int iL = 2 + blockIdx.x * TI, jL = 2 + blockIdx.y * T::TJ;
  const int k0 = 2 + blockIdx.z * planes;
  const int k1 = k0 + planes - 1 < d.nz + 1 ? k0 + planes - 1 : d.nz + 1;
  const int64_t sz = (int64_t)d.NX * d.NY;

  // own cells (i, j[r]) at slot (tx + 1, ty + 1 + 4r); loaded where inside
  // the array (a ghost cell is a neighbour), computed where interior
  const int i = iL + tx;
  int off[2], y[2], hoff[2];
  bool ld[2], cmp[2];
  Coef3 ay[2];
  for (int r = 0; r < 2; ++r) {
    const int j = jL + ty + R * r;
    y[r] = ty + 1 + R * r;
    ld[r] = i <= d.NX && j <= d.NY;
    cmp[r] = i <= d.nx + 1 && j <= d.ny + 1;
    off[r] = ld[r] ? (i - 1) + d.NX * (j - 1) : 0;
    hoff[r] = cmp[r] ? ((i - 1) >> 1) + NXh * (j - 1) : 0;
    ay[r] = cmp[r] ? coef_at(Ay, d.ny, j) : Coef3{0.0, 0.0, 0.0};
  }
  const Coef3 ax = i <= d.nx + 1 ? coef_at(Ax, d.nx, i) : Coef3{0.0, 0.0, 0.0};
  // halo cell of this thread: top row, bottom row, left and right columns
  int hx = 0, hy = 0;
  if (tid < kResPW) {
    hx = tid;
  } else if (tid < 2 * kResPW) {
    hx = tid - kResPW, hy = kResSH - 1;
  } else if (tid < kResHalo) {
    const int q = tid - 2 * kResPW;
    hx = (q & 1) ? kResPW - 1 : 0;
    hy = 1 + (q >> 1);
  }
  const int hi = iL - 1 + hx, hj = jL - 1 + hy;
  const bool hld = tid < kResHalo && hi >= 1 && hi <= d.NX && hj >= 1 && hj <= d.NY;
  const int hof = hld ? (hi - 1) + d.NX * (hj - 1) : 0;

  // raw loads of a plane (pc and e kept apart: the sum is formed when the
  // plane is stored, a step later)
  double rp[3];
  float re[3];
  auto load = [&](int k) {
    const int64_t base = (int64_t)(k - 1) * sz;
    for (int r = 0; r < 2; ++r) {
      rp[r] = ld[r] ? pc[base + off[r]] : 0.0;
      if (kUpdate) re[r] = ld[r] ? e[base + off[r]] : 0.0f;
    }
    rp[2] = hld ? pc[base + hof] : 0.0;
    if (kUpdate) re[2] = hld ? e[base + hof] : 0.0f;
  };
  auto val = [&](int q) -> double { return kUpdate ? rp[q] + (double)re[q] : rp[q]; };
  auto store = [&](int k, double *own) {
    double(*P)[kResPW] = pl[k & 1];
    for (int r = 0; r < 2; ++r) {
      own[r] = val(r);
      P[y[r]][tx + 1] = own[r];
    }
    if (tid < kResHalo) P[hy][hx] = val(2);
  };
  auto load_rhs = [&](int k, double *rh) {
    const int64_t base = (int64_t)NXh * d.NY * (k - 1);
    for (int r = 0; r < 2; ++r) {
      const bool red = ((i + jL + ty + R * r + k) & 1) != 0;
      rh[r] = cmp[r] ? (red ? bred : bblack)[base + hoff[r]] : not_fluid_marker();
    }
  };

  double vm[2], v0[2], vp[2], m = 0.0;
  // step k: rh holds the rhs of plane k, rhn receives those of plane k+1
  // (two buffers used alternately, so no register waits on a load in flight)
  auto step = [&](int k, const double *rh, double *rhn) {
    store(k + 1, vp);  // plane k+1 (its loads were issued a step ago)
    if (k + 1 <= k1) {
      load(k + 2);
      load_rhs(k + 1, rhn);
    }
    const double(*P)[kResPW] = pl[k & 1];
    const Coef3 az = coef_at(Az, d.nz, k);
    const int64_t base = (int64_t)(k - 1) * sz;
    const int64_t hbase = (int64_t)NXh * d.NY * (k - 1);
    for (int r = 0; r < 2; ++r) {
      if (!cmp[r]) continue;
      if (kUpdate) pc_out[base + off[r]] = v0[r];
      float rv = not_fluid_marker<float>();
      if (!sor_skip(rh[r])) {
        const int x = tx + 1, yy = y[r];
        const double res = (rh[r] - ax.m * P[yy][x - 1] - ax.p * P[yy][x + 1] - ay[r].m * P[yy - 1][x] -
                            ay[r].p * P[yy + 1][x] - az.m * vm[r] - az.p * vp[r]) -
                           (ax.c + ay[r].c + az.c) * v0[r];
        rv = (float)res;
        const double ar = fabs(res);
        m = m > ar ? m : ar;
      }
      if (kStore) {
        const bool red = ((i + jL + ty + R * r + k) & 1) != 0;
        (red ? rred : rblack)[hbase + hoff[r]] = rv;
      }
    }
    for (int r = 0; r < 2; ++r) {
      vm[r] = v0[r];
      v0[r] = vp[r];
    }
    __syncthreads();
  };

  double rha[2], rhb[2];
  load(k0 - 1);
  for (int r = 0; r < 2; ++r) vm[r] = val(r);
  load(k0);
  store(k0, v0);
  load(k0 + 1);
  load_rhs(k0, rha);
  __syncthreads();
  for (int k = k0; k <= k1; k += 2) {
    step(k, rha, rhb);
    if (k + 1 <= k1) step(k + 1, rhb, rha);
  }
  wg_max_store(m, partials);
}

// the FP32 marker in every entry of a colour-split residual array (the
// entries of non-interior cells are never written afterwards)
__global__ void k_mx_fill_marker(float *__restrict__ r, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) r[q] = not_fluid_marker<float>();
}

// after an odd number of updates the iterate is in pc2
__global__ void k_mx_finish(const SorState *__restrict__ st, double *__restrict__ pc, const double *__restrict__ pc2,
                            int64_t n) {
  if ((st->call_iters & 1) == 0) return;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    pc[q] = pc2[q];
}

// call set-up of the mixed solve (reset_iter: also nIterPcor and derr2)
__global__ void k_mx_reset(SorState *st, int reset_iter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (reset_iter) { st->n_iter = 0; st->derr2 = 0.0; st->derr_std_st = 0.0; st->exhausted = 0; }
    st->derr4 = 11111.0;
    st->done = 0;
    st->call_iters = 0;
  }
}

// end of a cycle: ||r||inf, nIterPcor = start + sweeps, updates so far, stop test
__global__ void __launch_bounds__(1024) k_mx_end(SorState *st, const double *partials, int64_t n, double tol,
                                                 long long start, long long sweeps, long long budget,
                                                 long long updates) {
  if (st->done) return;
  __shared__ double s_r;
  block_max_partials(partials, n, &s_r);
  if (threadIdx.x != 0) return;
  st->derr4 = s_r;
  st->derr2 = s_r;
  st->n_iter = start + sweeps;
  st->call_iters = updates;
  if (s_r < tol || sweeps >= budget) {
    st->done = 1;
    st->exhausted += !(s_r < tol);
  }
}

// ---------------------------------------------------------------- host

using Sweep32Launch = void (*)(dim3, dim3, hipStream_t, const FusedArgsT<float> &, int, int, const SorState *,
                               double *);
using Batch32Launch = void (*)(dim3, dim3, hipStream_t, const BatchBlockT<float> *, int, int);
struct Sweep32Variant {
  int ti, tj, nt, depth, minb, nit;
  Sweep32Launch launch, launch1;
  int family;  // 0: SorTile, 1, 2: SorPairTile (coefficients in registers, LDS)
  Batch32Launch batch, batch1;  // several blocks in one launch (pair-owned tiles only)
};

template <int TI, int TJ, int NT, int D, int MINB, int NIT>
static void launch_sweep32(dim3 g, dim3 blk, hipStream_t s, const FusedArgsT<float> &a, int gx, int gy,
                           const SorState *st, double *partials) {
  hipLaunchKernelGGL((k_sor_fused<float, TI, TJ, NT, D, NIT, MINB>), g, blk, 0, s, a, gx, gy, 0, st, partials);
}

template <int TI, int TJ, int NT, int MINW, int NIT, bool CREG>
static void launch_pair32(dim3 g, dim3 blk, hipStream_t s, const FusedArgsT<float> &a, int gx, int gy,
                          const SorState *st, double *partials) {
  hipLaunchKernelGGL((k_sor_pair<float, TI, TJ, NT, NIT, MINW, CREG>), g, blk, 0, s, a, gx, gy, 0, st, partials);
}

template <int TI, int TJ, int NT, int MINW, int NIT, bool CREG>
static void launch_pair32_batch(dim3 g, dim3 blk, hipStream_t s, const BatchBlockT<float> *tab, int nblk, int parity) {
  hipLaunchKernelGGL((k_sor_pair_batch<float, TI, TJ, NT, NIT, MINW, CREG>), g, blk, 0, s, tab, nblk, parity, 0,
                     nullptr);
}

// SV(tile_i, tile_j, threads, prefetch depth, waves per SIMD the registers
// must allow, sweeps per pass): SorTile; PS(tile_i, tile_j, threads, waves
// per SIMD, sweeps per pass, coefficients in registers): SorPairTile
#define SV(TI, TJ, NT, D, MINB, NIT) \
  { TI, TJ, NT, D, MINB, NIT, launch_sweep32<TI, TJ, NT, D, MINB, NIT>, launch_sweep32<TI, TJ, NT, D, MINB, 1>, 0, \
    nullptr, nullptr }
#define PS(TI, TJ, NT, MINW, NIT, CREG) \
  { TI, TJ, NT, 1, MINW, NIT, launch_pair32<TI, TJ, NT, MINW, NIT, CREG>, launch_pair32<TI, TJ, NT, MINW, 1, CREG>, \
    CREG ? 1 : 2, launch_pair32_batch<TI, TJ, NT, MINW, NIT, CREG>, launch_pair32_batch<TI, TJ, NT, MINW, 1, CREG> }
static const Sweep32Variant kSweep32[] = {
    SV(64, 16, 256, 1, 1, 1),
    SV(64, 32, 512, 1, 1, 1),
    SV(64, 32, 512, 2, 1, 1),
    SV(128, 16, 512, 1, 1, 1),
    SV(128, 16, 512, 2, 1, 1),
    SV(64, 32, 512, 1, 3, 1),
    SV(64, 16, 256, 1, 1, 2),
    SV(64, 32, 512, 1, 1, 2),
    SV(128, 16, 512, 1, 1, 2),
    SV(128, 32, 512, 1, 1, 2),
    SV(64, 16, 512, 1, 1, 2),
    SV(64, 32, 1024, 1, 1, 1),
    SV(64, 32, 1024, 1, 1, 2),
    SV(128, 32, 1024, 1, 1, 1),
    SV(128, 16, 1024, 1, 1, 1),
    PS(64, 32, 512, 1, 1, 1),
    PS(64, 32, 512, 1, 2, 1),
    PS(64, 28, 512, 4, 1, 1),
    PS(64, 28, 512, 4, 2, 0),
    PS(64, 32, 512, 4, 2, 0),
    PS(128, 16, 512, 4, 1, 1),
    PS(64, 32, 1024, 1, 2, 1),
    PS(64, 32, 1024, 1, 1, 1),
    PS(64, 22, 1024, 1, 2, 1),
    PS(64, 22, 512, 1, 2, 1),
    PS(64, 32, 1024, 1, 2, 0),
    PS(64, 32, 768, 1, 2, 1),
};
#undef SV
#undef PS
constexpr int kNumSweep32 = (int)(sizeof(kSweep32) / sizeof(kSweep32[0]));

struct MixedConfig {
  int precision = 0;        // 0: FP64 method, 1: mixed
  int variant = 26;         // the fastest at 256^3 on MI355X (pair-owned 64x32, 768 threads, 2 sweeps per pass)
  int planes = 32;
  int sweeps = 128;         // FP32 sweeps per correction (u, v, w within 4.2e-7 of FP64 after 100 steps at 256^3)
  double tol = 1e-8;        // max-norm of the FP64 residual
};
static MixedConfig g_mx;
static SorState *g_mx_probe = nullptr;

int precision_mode() { return g_mx.precision; }

int64_t residual_workgroups(const Block &b);
int ensure_split(Block &b);

static int ensure_mixed(Block &b) {
  const int NXh = (b.d.NX + 1) / 2;
  const int64_t nh = (int64_t)NXh * b.d.NY * b.d.NZ;
  if (b.mx_n != b.n2) {
    for (int q = 0; q < 2; ++q) {
      if (dev_alloc(&b.e32[q], b.n2) || dev_alloc(&b.r32[q], nh)) return -1;
      // ghost cells of the corrections and residuals are never written: set once
      BCFD_HIP(hipMemsetAsync(b.e32[q], 0, sizeof(float) * (size_t)b.n2, rt().stream));
      hipLaunchKernelGGL(k_mx_fill_marker, dim3(div_up(nh, 256)), dim3(256), 0, rt().stream, b.r32[q], nh);
      BCFD_CHECK_LAUNCH();
    }
    b.mx_n = b.n2;
  }
  if (b.pc2_n != b.n2) {
    if (dev_alloc(&b.pc2, b.n2)) return -1;
    b.pc2_n = b.n2;
  }
  if (ensure_split(b)) return -1;
  if (grow(&b.partials, &b.cap_partials, residual_workgroups(b))) return -1;
  return 0;
}

// the residual launch of a block: 256 x 8 tiles of 32 planes, 64 x 32 tiles
// where the block is at most 128 cells wide, and fewer planes per workgroup
// (down to 2) while the launch would have fewer workgroups than CUs.  Every
// shape gives the same values and the same maximum.
struct ResShape {
  int ti, planes;
  dim3 grid;
};
static ResShape res_shape(const Dims &d) {
  ResShape r;
  r.ti = d.nx > 128 ? 256 : 64;
  const int tj = 2 * (kResNT / r.ti);
  r.planes = kResPlanes;
  r.grid = dim3(div_up(d.nx, r.ti), div_up(d.ny, tj), div_up(d.nz, r.planes));
  while (r.planes > 2 && (int64_t)r.grid.x * r.grid.y * r.grid.z < 256) {
    r.planes /= 2;
    r.grid.z = div_up(d.nz, r.planes);
  }
  return r;
}

int64_t residual_workgroups(const Block &b) {
  const ResShape r = res_shape(b.d);
  return (int64_t)r.grid.x * r.grid.y * r.grid.z;
}

template <bool kUpdate, bool kStore>
static int64_t launch_resupd(Block &b, const double *src, const float *e, double *dst, float *rred, float *rblack,
                             const SorState *st) {
  const int NXh = (b.d.NX + 1) / 2;
  const ResShape r = res_shape(b.d);
  if (r.ti == 256)
    hipLaunchKernelGGL((k_mx_resupd<256, kUpdate, kStore>), r.grid, dim3(kResNT), 0, rt().stream, b.d, NXh, r.planes,
                       src, e, dst, b.bs[1], b.bs[0], b.A[0], b.A[1], b.A[2], rred, rblack, st, b.partials);
  else
    hipLaunchKernelGGL((k_mx_resupd<64, kUpdate, kStore>), r.grid, dim3(kResNT), 0, rt().stream, b.d, NXh, r.planes,
                       src, e, dst, b.bs[1], b.bs[0], b.A[0], b.A[1], b.A[2], rred, rblack, st, b.partials);
  return (int64_t)r.grid.x * r.grid.y * r.grid.z;
}

// b/dt of the fluid cells into the colour-split FP64 arrays (marker elsewhere)
int split_bdt(Block &b, double dt);

// residual of pc (src, with e added when e != nullptr, written to dst)
static int64_t residual_launch(Block &b, const double *src, const float *e, double *dst, const SorState *st) {
  if (e) return launch_resupd<true, true>(b, src, e, dst, b.r32[1], b.r32[0], st);
  return launch_resupd<false, true>(b, src, e, dst, b.r32[1], b.r32[0], st);
}

// k sweeps from a zero correction, in passes of up to `nit` sweeps; returns
// the array holding the result
static float *sweeps_launch(Block &b, int variant, int planes, int k, const SorState *st) {
  const Sweep32Variant &v = kSweep32[variant];
  const int gx = (int)div_up(b.d.nx, v.ti), gy = (int)div_up(b.d.ny, v.tj), gz = (int)div_up(b.d.nz, planes);
  FusedArgsT<float> a;
  a.d = b.d;
  a.NXh = (b.d.NX + 1) / 2;
  a.KS = planes;
  a.bred = b.r32[1];
  a.bblack = b.r32[0];
  a.Ax = b.A[0];
  a.Ay = b.A[1];
  a.Az = b.A[2];
  a.omega = b.omega;
  float *out = b.e32[0];
  int q = 0;
  for (int done = 0; done < k; ++q) {
    const int L = k - done < v.nit ? k - done : v.nit;
    a.zero_src = q == 0;
    a.src = (q & 1) ? b.e32[0] : b.e32[1];
    a.dst = (q & 1) ? b.e32[1] : b.e32[0];
    out = a.dst;
    (L == v.nit ? v.launch : v.launch1)(dim3((unsigned)(gx * gy * gz)), dim3(v.nt), rt().stream, a, gx, gy, st,
                                         b.partials);
    done += L;
  }
  return out;
}

int sor_timing_on();
int sor_poll(const SorState *st, long long &k, bool &stop);
int sor_timing_begin(hipEvent_t *e0);
int sor_timing_end(hipEvent_t e0, const Block &b, long long start);
int sor_timing_end_set(hipEvent_t e0, const std::vector<Block *> &blocks, const std::vector<long long> &start);
int set_poll(const std::vector<Block *> &blocks, long long &k, bool &stop);
int read_states(const std::vector<Block *> &blocks, std::vector<SorState> &out);

int mixed_sor(Block &b, int64_t pc_ita_max, double dt) {
  hipStream_t s = rt().stream;
  if (ensure_mixed(b) || split_bdt(b, dt)) return -1;
  BCFD_HIP(hipMemcpyAsync(b.pc2, b.f[BF_PC], sizeof(double) * (size_t)b.n2, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_mx_reset, dim3(1), dim3(1), 0, s, b.sor, 0);
  BCFD_CHECK_LAUNCH();
  SorState *h = (SorState *)rt().h_pinned;
  BCFD_HIP(hipMemcpyAsync(h, b.sor, sizeof(SorState), hipMemcpyDeviceToHost, s));
  BCFD_HIP(hipStreamSynchronize(s));
  const long long start = h->n_iter;
  long long budget = pc_ita_max + 1 - start;
  if (budget < 1) budget = 1;
  hipEvent_t e0 = nullptr;
  if (sor_timing_begin(&e0)) return -1;
  const int k = g_mx.sweeps;
  const long long cycles = (budget + k - 1) / k;
  int64_t nwg = residual_launch(b, b.f[BF_PC], nullptr, nullptr, b.sor);
  // the residual of the call's starting pc (the first correction's right-hand
  // side) never stops the call: at least one correction runs, as the
  // reference's loop runs at least one iteration (PcorVcor.F90:309)
  hipLaunchKernelGGL(k_mx_end, dim3(1), dim3(1024), 0, s, b.sor, b.partials, nwg, -1.0, start, 0ll, budget, 0ll);
  BCFD_CHECK_LAUNCH();
  const long long chunk = (48 + k - 1) / k;  // host polls about every 48 sweeps
  long long polls = 0;
  for (long long c = 0; c < cycles; ++c) {
    // the last correction takes only the sweeps left in the budget
    const int kc = (int)((budget - c * k) < k ? (budget - c * k) : k);
    const float *e = sweeps_launch(b, g_mx.variant, g_mx.planes, kc, b.sor);
    // update c reads the iterate from pc (c even) or pc2 and writes the other
    const double *src = (c & 1) ? b.pc2 : b.f[BF_PC];
    double *dst = (c & 1) ? b.f[BF_PC] : b.pc2;
    nwg = residual_launch(b, src, e, dst, b.sor);
    const long long done_sweeps = c * k + kc;
    hipLaunchKernelGGL(k_mx_end, dim3(1), dim3(1024), 0, s, b.sor, b.partials, nwg, g_mx.tol, start, done_sweeps,
                       budget, c + 1);
    BCFD_CHECK_LAUNCH();
    if ((c + 1) % chunk == 0 && c + 1 < cycles) {
      bool stop = false;
      if (sor_poll(b.sor, polls, stop)) return -1;
      if (stop) break;
    }
  }
  if (sor_timing_end(e0, b, start)) return -1;
  hipLaunchKernelGGL(k_mx_finish, dim3(1024), dim3(256), 0, s, b.sor, b.f[BF_PC], b.pc2, b.n2);
  BCFD_CHECK_LAUNCH();
  return 0;
}

// ------------------------------------------------ several blocks at once
// The mixed solves of a set of blocks (poissonSolver's fine blocks,
// PcorVcor.F90:86-129): blocks whose calls start at the same iteration share
// one schedule of corrections, and the FP32 sweeps of a pass run for all of
// them in one launch of the batched pair-owned kernel (a block whose call
// has stopped skips its tiles); the residual updates and the end-of-correction
// bookkeeping stay per block.  Each block sees the kernels, arguments and
// order of mixed_sor: the same bits.
static BatchBlockT<float> *g_mx_tab = nullptr;
static int64_t g_mx_tab_cap = 0;
bool mixed_batchable() { return kSweep32[g_mx.variant].batch != nullptr; }

// one call on each of `blocks`, all at start iteration `start`
static int mixed_group(const std::vector<Block *> &blocks, long long start, int64_t pc_ita_max) {
  hipStream_t s = rt().stream;
  const Sweep32Variant &v = kSweep32[g_mx.variant];
  const int nb = (int)blocks.size();
  long long budget = pc_ita_max + 1 - start;
  if (budget < 1) budget = 1;
  const int k = g_mx.sweeps;
  const long long cycles = (budget + k - 1) / k;
  // entries [0, nb): the first pass of a correction (zero source, into
  // e32[0]); [nb, 2 nb): later passes, parity 0 e32[1] -> e32[0], parity 1
  // e32[0] -> e32[1]
  std::vector<BatchBlockT<float>> tab((size_t)(2 * nb));
  int64_t total = 0;
  for (int i = 0; i < nb; ++i) {
    Block &b = *blocks[i];
    const int gx = (int)div_up(b.d.nx, v.ti), gy = (int)div_up(b.d.ny, v.tj), gz = (int)div_up(b.d.nz, g_mx.planes);
    for (int t = 0; t < 2; ++t) {
      BatchBlockT<float> &e = tab[(size_t)(t * nb + i)];
      for (int par = 0; par < 2; ++par) {
        FusedArgsT<float> &a = e.a[par];
        a.d = b.d;
        a.NXh = (b.d.NX + 1) / 2;
        a.KS = g_mx.planes;
        a.bred = b.r32[1];
        a.bblack = b.r32[0];
        a.Ax = b.A[0];
        a.Ay = b.A[1];
        a.Az = b.A[2];
        a.omega = b.omega;
        a.zero_src = t == 0 && par == 0;
        a.src = par ? b.e32[0] : b.e32[1];
        a.dst = par ? b.e32[1] : b.e32[0];
      }
      e.gx = gx;
      e.gy = gy;
      e.ntiles = gx * gy * gz;
      e.first = (int)total;
      e.st = b.sor;
      e.start = start;
    }
    total += (int64_t)gx * gy * gz;
  }
  if (grow(&g_mx_tab, &g_mx_tab_cap, 2 * nb)) return -1;
  BCFD_HIP(hipMemcpyAsync(g_mx_tab, tab.data(), sizeof(BatchBlockT<float>) * 2 * nb, hipMemcpyHostToDevice, s));
  // the host table is freed on return: wait for the copy
  BCFD_HIP(hipStreamSynchronize(s));
  hipEvent_t e0 = nullptr;
  if (sor_timing_begin(&e0)) return -1;
  std::vector<int64_t> nwg((size_t)nb);
  for (int i = 0; i < nb; ++i) {
    Block &b = *blocks[i];
    nwg[i] = residual_launch(b, b.f[BF_PC], nullptr, nullptr, b.sor);
    hipLaunchKernelGGL(k_mx_end, dim3(1), dim3(1024), 0, s, b.sor, b.partials, nwg[i], -1.0, start, 0ll, budget, 0ll);
  }
  BCFD_CHECK_LAUNCH();
  const long long chunk = (48 + k - 1) / k;
  long long polls = 0;
  for (long long c = 0; c < cycles; ++c) {
    const int kc = (int)((budget - c * k) < k ? (budget - c * k) : k);
    int q = 0;
    for (int done = 0; done < kc; ++q) {
      const int L = kc - done < v.nit ? kc - done : v.nit;
      const BatchBlockT<float> *t = q == 0 ? g_mx_tab : g_mx_tab + nb;
      (L == v.nit ? v.batch : v.batch1)(dim3((unsigned)total), dim3(v.nt), s, t, nb, q & 1);
      done += L;
    }
    const bool in0 = ((q - 1) & 1) == 0;   // the last pass wrote e32[0]
    for (int i = 0; i < nb; ++i) {
      Block &b = *blocks[i];
      const float *e = in0 ? b.e32[0] : b.e32[1];
      const double *src = (c & 1) ? b.pc2 : b.f[BF_PC];
      double *dst = (c & 1) ? b.f[BF_PC] : b.pc2;
      nwg[i] = residual_launch(b, src, e, dst, b.sor);
      hipLaunchKernelGGL(k_mx_end, dim3(1), dim3(1024), 0, s, b.sor, b.partials, nwg[i], g_mx.tol, start,
                         c * k + kc, budget, c + 1);
    }
    BCFD_CHECK_LAUNCH();
    if ((c + 1) % chunk == 0 && c + 1 < cycles) {
      bool stop = false;
      if (set_poll(blocks, polls, stop)) return -1;
      if (stop) break;
    }
  }
  return sor_timing_end_set(e0, blocks, std::vector<long long>(blocks.size(), start));
}

int mixed_sor_set(const std::vector<Block *> &blocks, int64_t pc_ita_max, double dt) {
  hipStream_t s = rt().stream;
  for (Block *b : blocks) {
    if (ensure_mixed(*b) || split_bdt(*b, dt)) return -1;
    BCFD_HIP(hipMemcpyAsync(b->pc2, b->f[BF_PC], sizeof(double) * (size_t)b->n2, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_mx_reset, dim3(1), dim3(1), 0, s, b->sor, 0);
  }
  BCFD_CHECK_LAUNCH();
  std::vector<SorState> st0;
  if (read_states(blocks, st0)) return -1;
  std::vector<long long> start(blocks.size());
  for (size_t i = 0; i < blocks.size(); ++i) start[i] = st0[i].n_iter;
  std::vector<bool> taken(blocks.size(), false);
  for (size_t i = 0; i < blocks.size(); ++i) {
    if (taken[i]) continue;
    std::vector<Block *> group;
    for (size_t j = i; j < blocks.size(); ++j)
      if (!taken[j] && start[j] == start[i]) {
        group.push_back(blocks[j]);
        taken[j] = true;
      }
    if (mixed_group(group, start[i], pc_ita_max)) return -1;
  }
  for (Block *b : blocks)
    hipLaunchKernelGGL(k_mx_finish, dim3(1024), dim3(256), 0, s, b->sor, b->f[BF_PC], b->pc2, b->n2);
  BCFD_CHECK_LAUNCH();
  return 0;
}

// `iters` FP32 sweeps of a variant on the residual of the current pc, timed
int mixed_time(Block &b, int variant, int planes, int iters, double dt, double *ms) {
  hipStream_t s = rt().stream;
  if (ensure_mixed(b) || split_bdt(b, dt)) return -1;
  if (!g_mx_probe && dev_alloc(&g_mx_probe, 1)) return -1;
  hipLaunchKernelGGL(k_mx_reset, dim3(1), dim3(1), 0, s, g_mx_probe, 1);
  residual_launch(b, b.f[BF_PC], nullptr, nullptr, g_mx_probe);
  hipEvent_t e0, e1;
  BCFD_HIP(hipEventCreate(&e0));
  BCFD_HIP(hipEventCreate(&e1));
  BCFD_HIP(hipEventRecord(e0, s));
  sweeps_launch(b, variant, planes, iters, g_mx_probe);
  BCFD_HIP(hipEventRecord(e1, s));
  BCFD_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  BCFD_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  BCFD_CHECK_LAUNCH();
  *ms = t;
  return 0;
}

// max |b/dt - A pc| over the fluid cells of `pc`, per workgroup into the
// block's partials (the FP64 method's residual stopping test; the caller
// reduces them); returns the number of workgroups
int64_t residual_check_launch(Block &b, const double *pc, const SorState *st) {
  return launch_resupd<false, false>(b, pc, nullptr, nullptr, nullptr, nullptr, st);
}

// max-norm of b/dt - A pc over the fluid cells (diagnostic, FP64)
int residual_norm(Block &b, double dt, double *out) {
  hipStream_t s = rt().stream;
  if (ensure_mixed(b) || split_bdt(b, dt)) return -1;
  if (!g_mx_probe && dev_alloc(&g_mx_probe, 1)) return -1;
  hipLaunchKernelGGL(k_mx_reset, dim3(1), dim3(1), 0, s, g_mx_probe, 1);
  const int64_t nwg = residual_launch(b, b.f[BF_PC], nullptr, nullptr, g_mx_probe);
  hipLaunchKernelGGL(k_mx_end, dim3(1), dim3(1024), 0, s, g_mx_probe, b.partials, nwg, -1.0, 0ll, 0ll, 1ll, 0ll);
  BCFD_CHECK_LAUNCH();
  SorState *h = (SorState *)rt().h_pinned;
  BCFD_HIP(hipMemcpyAsync(h, g_mx_probe, sizeof(SorState), hipMemcpyDeviceToHost, s));
  BCFD_HIP(hipStreamSynchronize(s));
  *out = h->derr4;
  return 0;
}

}  // namespace bcfd

using namespace bcfd;

extern "C" {

int biocfd_set_precision(int mode) {
  BCFD_API_LOCK();
  if (mode != 0 && mode != 1) {
    set_error("set_precision: mode %d is neither 0 (FP64) nor 1 (mixed)", mode);
    return -1;
  }
  g_mx.precision = mode;
  return 0;
}

int biocfd_mixed_config(int variant, int planes_per_workgroup, int sweeps_per_correction, double residual_tol) {
  BCFD_API_LOCK();
  if (variant < 0 || variant >= kNumSweep32) {
    set_error("mixed_config: variant %d outside 0..%d", variant, kNumSweep32 - 1);
    return -1;
  }
  if (planes_per_workgroup < 1 || planes_per_workgroup > 4096 || sweeps_per_correction < 1 ||
      sweeps_per_correction > 1024) {
    set_error("mixed_config: planes %d / sweeps %d out of range", planes_per_workgroup, sweeps_per_correction);
    return -1;
  }
  g_mx.variant = variant;
  g_mx.planes = planes_per_workgroup;
  g_mx.sweeps = sweeps_per_correction;
  g_mx.tol = residual_tol;
  return 0;
}

int biocfd_mixed_variant_info(int variant, int info[7]) {
  BCFD_API_LOCK();
  if (variant < 0 || variant >= kNumSweep32) {
    set_error("mixed_variant_info: variant %d outside 0..%d", variant, kNumSweep32 - 1);
    return -1;
  }
  const Sweep32Variant &v = kSweep32[variant];
  info[0] = v.ti;
  info[1] = v.tj;
  info[2] = v.nt;
  info[3] = v.depth;
  info[4] = v.minb;
  info[5] = v.nit;
  info[6] = v.family;
  return 0;
}

int biocfd_mixed_time(int id, int variant, int planes_per_workgroup, int iters, double dt, double *ms) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (!b->A[0]) { set_error("coefficientMatrix has not been called"); return -1; }
  if (variant < 0 || variant >= kNumSweep32 || planes_per_workgroup < 1 || planes_per_workgroup > 4096 ||
      iters < 1) {
    set_error("mixed_time: bad variant %d / planes %d / iterations %d", variant, planes_per_workgroup, iters);
    return -1;
  }
  return mixed_time(*b, variant, planes_per_workgroup, iters, dt, ms);
}

int biocfd_residual_norm(int id, double dt, double *out) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (!b->A[0]) { set_error("coefficientMatrix has not been called"); return -1; }
  return residual_norm(*b, dt, out);
}

}  // extern "C"
