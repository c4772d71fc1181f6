// Failure detection around the pressure solve.  The reference has none: a
// field that turns non-finite keeps stepping, and a red-black SOR call that
// reaches pcItaMax ends silently, its iteration count printed and nothing
// more (PcorVcor.F90:181).  The checks here are opt-in, so the default
// behaviour stays the reference's:
//
//  * a scan of u, v, w and p of a block (all cells, ghost layer included):
//    the count of non-finite values, the first one's field and (i,j,k), and
//    the largest |velocity component| among the finite values;
//  * the count of SOR calls of the last poissonSolver call that ended on
//    their iteration budget rather than their tolerance (kept in the SOR
//    state by the end-of-call kernels);
//  * a switch that makes biocfd_poisson_solver run both checks on its blocks
//    once it has finished and return -3 (non-finite value or velocity over a
//    limit) or -4 (budget exhausted) with the block named in the error text.
//
// The scan reads 32 B per cell (four FP64 fields) once: one HBM-bound pass,
// about a tenth of a millisecond for a 258^3 block.
#include "biocfd_rt.hpp"

#include <cmath>

namespace bcfd {
namespace {

struct HealthOut {
  unsigned long long nonfinite;
  unsigned long long vmax_bits;   // bit pattern of a double >= 0: ordered as unsigned
  unsigned long long first_bad;   // field * n + cell of the first non-finite value, or ~0
  unsigned long long pad;
};

struct HealthConfig {
  bool after_solve = false;
  double vel_limit = 0.0;         // > 0: a finite |velocity component| above it fails
  bool exhausted_fails = false;
};

HealthConfig g_cfg;
HealthOut *g_out = nullptr;       // device: one slot per block
int g_out_device = -1;            // the device g_out lives on
HealthOut *g_host = nullptr;      // pinned: the same slots, read back together
int *g_host_ex = nullptr;         // pinned: SorState::exhausted per block

constexpr int kHealthThreads = 256;
constexpr int kHealthMaxGroups = 2048;

__global__ void k_health_clear(HealthOut *o) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    o->nonfinite = 0ull;
    o->vmax_bits = 0ull;
    o->first_bad = ~0ull;
    o->pad = 0ull;
  }
}

__global__ void __launch_bounds__(kHealthThreads) k_health(const double *__restrict__ u,
                                                           const double *__restrict__ v,
                                                           const double *__restrict__ w,
                                                           const double *__restrict__ p, int64_t n,
                                                           HealthOut *__restrict__ o) {
  unsigned long long bad = 0ull, first = ~0ull;
  double vmax = 0.0;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const double a[4] = {u[q], v[q], w[q], p[q]};
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      if (!__builtin_isfinite(a[f])) {
        ++bad;
        const unsigned long long at = (unsigned long long)f * (unsigned long long)n + (unsigned long long)q;
        first = at < first ? at : first;
      } else if (f < 3) {
        const double m = fabs(a[f]);
        vmax = m > vmax ? m : vmax;
      }
    }
  }
  for (int s = 32; s > 0; s >>= 1) {
    bad += __shfl_xor(bad, s, 64);
    const unsigned long long f2 = __shfl_xor(first, s, 64);
    first = f2 < first ? f2 : first;
    const double v2 = __shfl_xor(vmax, s, 64);
    vmax = v2 > vmax ? v2 : vmax;
  }
  __shared__ unsigned long long s_bad[kHealthThreads / 64], s_first[kHealthThreads / 64];
  __shared__ double s_v[kHealthThreads / 64];
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { s_bad[wv] = bad; s_first[wv] = first; s_v[wv] = vmax; }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int q = 1; q < kHealthThreads / 64; ++q) {
    bad += s_bad[q];
    first = s_first[q] < first ? s_first[q] : first;
    vmax = s_v[q] > vmax ? s_v[q] : vmax;
  }
  if (bad) {
    atomicAdd(&o->nonfinite, bad);
    atomicMin(&o->first_bad, first);
  }
  atomicMax(&o->vmax_bits, (unsigned long long)__double_as_longlong(vmax));
}

int ensure_buffers() {
  if (g_out && g_out_device != rt().device) g_out = nullptr;   // left on another device by an earlier init
  if (!g_out) {
    if (dev_alloc(&g_out, kMaxBlocks + 1)) return -1;
    g_out_device = rt().device;
  }
  if (!g_host) BCFD_HIP(hipHostMalloc((void **)&g_host, sizeof(HealthOut) * (kMaxBlocks + 1), hipHostMallocDefault));
  if (!g_host_ex) BCFD_HIP(hipHostMalloc((void **)&g_host_ex, sizeof(int) * (kMaxBlocks + 1), hipHostMallocDefault));
  return 0;
}

// queue the scan of block g into slot g and the read-back of the slot and of
// the block's exhaustion count (no wait)
int enqueue_scan(int g) {
  Block &b = rt().blk[g];
  hipStream_t s = rt().stream;
  HealthOut *o = g_out + g;
  hipLaunchKernelGGL(k_health_clear, dim3(1), dim3(1), 0, s, o);
  const int64_t groups = std::min<int64_t>(kHealthMaxGroups, (b.n2 + kHealthThreads - 1) / kHealthThreads);
  hipLaunchKernelGGL(k_health, dim3((unsigned)groups), dim3(kHealthThreads), 0, s, b.f[BF_U], b.f[BF_V],
                     b.f[BF_W], b.f[BF_P], b.n2, o);
  BCFD_CHECK_LAUNCH();
  BCFD_HIP(hipMemcpyAsync(g_host + g, o, sizeof(HealthOut), hipMemcpyDeviceToHost, s));
  BCFD_HIP(hipMemcpyAsync(g_host_ex + g, &b.sor->exhausted, sizeof(int), hipMemcpyDeviceToHost, s));
  return 0;
}

// field name and 1-based (i,j,k) of a cell of the n2 layout
void locate(const Block &b, unsigned long long at, int *field, int64_t ijk[3]) {
  const int64_t n = b.n2;
  *field = (int)(at / (unsigned long long)n);
  const int64_t q = (int64_t)(at % (unsigned long long)n);
  ijk[0] = q % b.d.NX + 1;
  ijk[1] = (q / b.d.NX) % b.d.NY + 1;
  ijk[2] = q / ((int64_t)b.d.NX * b.d.NY) + 1;
}

const char *kFieldName[4] = {"u", "v", "w", "p"};

}  // namespace

// the checks biocfd_health_config switched on, over blocks first..last, after
// a pressure solve (0 when they are off or pass)
int health_after_solve(int first, int last) {
  if (!g_cfg.after_solve) return 0;
  if (ensure_buffers()) return -1;
  for (int g = first; g <= last; ++g)
    if (enqueue_scan(g)) return -1;
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  for (int g = first; g <= last; ++g) {
    const HealthOut &h = g_host[g];
    const Block &b = rt().blk[g];
    if (h.nonfinite) {
      int f;
      int64_t ijk[3];
      locate(b, h.first_bad, &f, ijk);
      set_error("block %d: %llu non-finite values in u, v, w, p after the pressure solve; first %s(%lld,%lld,%lld)",
                g, h.nonfinite, kFieldName[f], (long long)ijk[0], (long long)ijk[1], (long long)ijk[2]);
      return -3;
    }
    const double vmax = __builtin_bit_cast(double, h.vmax_bits);
    if (g_cfg.vel_limit > 0.0 && vmax > g_cfg.vel_limit) {
      set_error("block %d: velocity component %.6e over the limit %.6e after the pressure solve", g, vmax,
                g_cfg.vel_limit);
      return -3;
    }
  }
  if (g_cfg.exhausted_fails)
    for (int g = first; g <= last; ++g)
      if (g_host_ex[g]) {
        set_error("block %d: %d red-black SOR call(s) of the pressure solve reached pcItaMax before the tolerance", g,
                  g_host_ex[g]);
        return -4;
      }
  return 0;
}

}  // namespace bcfd

using namespace bcfd;

extern "C" {

int biocfd_health_config(int check_after_solve, double vel_limit, int exhausted_is_error) {
  BCFD_API_LOCK();
  if (!(vel_limit >= 0.0) || std::isinf(vel_limit)) {
    set_error("health_config: velocity limit %g (0 = no limit, else finite and > 0)", vel_limit);
    return -1;
  }
  g_cfg.after_solve = check_after_solve != 0;
  g_cfg.vel_limit = vel_limit;
  g_cfg.exhausted_fails = exhausted_is_error != 0;
  return 0;
}

int biocfd_health_check(int blk, int64_t *n_nonfinite, double *vel_max, int64_t where[4]) {
  BCFD_API_LOCK();
  Block *b = get_block(blk);
  if (!b) return -1;
  if (ensure_buffers() || enqueue_scan(blk)) return -1;
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  const HealthOut &h = g_host[blk];
  if (n_nonfinite) *n_nonfinite = (int64_t)h.nonfinite;
  if (vel_max) *vel_max = __builtin_bit_cast(double, h.vmax_bits);
  if (where) {
    where[0] = where[1] = where[2] = where[3] = -1;
    if (h.nonfinite) {
      int f;
      locate(*b, h.first_bad, &f, where + 1);
      where[0] = f;
    }
  }
  return 0;
}

int biocfd_sor_exhausted(int blk, int *count) {
  BCFD_API_LOCK();
  Block *b = get_block(blk);
  if (!b) return -1;
  if (!b->sor) { set_error("sor_exhausted: block %d has no SOR state", blk); return -1; }
  int c = 0;
  BCFD_HIP(hipMemcpyAsync(&c, &b->sor->exhausted, sizeof(int), hipMemcpyDeviceToHost, rt().stream));
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  *count = c;
  return 0;
}

}  // extern "C"
