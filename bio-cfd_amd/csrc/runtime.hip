// Runtime: device selection, block storage, host<->device transfers and the
// host-side set-up numerics (coefficient tables, body kinematics).
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>

#include "biocfd_rt.hpp"

namespace bcfd {

static thread_local char g_err[1024] = "";
static Runtime g_rt;
static std::mutex g_mu;
int g_alloc_fail_in = 0;

std::recursive_mutex &api_mutex() {
  static std::recursive_mutex m;
  return m;
}

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

Runtime &rt() { return g_rt; }

int ensure_init() {
  if (!g_rt.init) {
    set_error("biocfd_init has not been called");
    return -1;
  }
  return 0;
}

Block *get_block(int id) {
  if (ensure_init()) return nullptr;
  if (id < 1 || id > kMaxBlocks || !g_rt.blk[id].live) {
    set_error("block %d is not registered", id);
    return nullptr;
  }
  return &g_rt.blk[id];
}

// ---- narrow/widen tags on the host ----------------------------------------
static int64_t tag_size(const Block &b, int tag) {
  return (tag == BT_CELL || tag == BT_CELL_N) ? b.n2 : b.n3;
}

// ---- coefficient tables (host; results uploaded) --------------------------
// 3-point non-uniform Laplacian rows A(n,1..3) from cell-centre spacing,
// with the coarse block's boundary folds (reference CoefficientMatrix.f90:10-71).
static void laplace_rows(const std::vector<double> &cc, int64_t n, std::vector<double> &A) {
  A.assign((size_t)(3 * n), 0.0);
  for (int64_t j = 2; j <= n + 1; ++j) {
    const double h1 = cc[j - 1] - cc[j - 2];
    const double h2 = cc[j] - cc[j - 1];
    const double hs = h1 + h2;
    A[(j - 2) + 0 * n] = 2.0 / (h1 * hs);
    A[(j - 2) + 1 * n] = -2.0 / (h1 * h2);
    A[(j - 2) + 2 * n] = 2.0 / (h2 * hs);
  }
}

static void fold(std::vector<double> &A, int64_t n, bool dirichlet) {
  // first row folds its west coefficient, last row its east one
  if (dirichlet) {
    A[0 + n] = A[0 + n] - A[0];
    A[(n - 1) + n] = A[(n - 1) + n] - A[(n - 1) + 2 * n];
  } else {
    A[0 + n] = A[0 + n] + A[0];
    A[(n - 1) + n] = A[(n - 1) + n] + A[(n - 1) + 2 * n];
  }
  A[0] = 0.0;
  A[(n - 1) + 2 * n] = 0.0;
}

// Integer powers as the Fortran compiler expands x**2, x**3, x**4.
static inline double p2(double x) { return x * x; }
static inline double p3(double x) { return x * x * x; }
static inline double p4(double x) { const double s = x * x; return s * s; }

// One column of the 3rd-order non-uniform upwind tables from the three
// spacing ratios th (reference NavierStokes.f90:915-996 and :47-61).
static void upwind_column(const double th[3], double ca[6], double ck[6]) {
  const double a = th[0], t2 = th[1], c = th[2];
  double f[9];
  f[1] = c;
  f[2] = 2.0 * c + p2(c);
  f[3] = 3.0 * c + 3.0 * p2(c) + p3(c);
  f[4] = 4.0 * c + 6.0 * p2(c) + 4.0 * p3(c) + p4(c);
  f[5] = 1.0 / (a * t2);
  f[6] = (2.0 * a + 1.0) / (p2(a) * p2(t2));
  f[7] = (3.0 * p2(a) + 3.0 * a + 1.0) / (p3(a) * p3(t2));
  f[8] = (4.0 * p3(a) + 6.0 * p2(a) + 4.0 * a + 1.0) / (p4(a) * p4(t2));
  // rows 5..10 of the elimination table (real-exponent powers via pow)
  const double s51 = -1.0 / p4(t2), s52 = f[4], s53 = f[4], s54 = s51;
  const double s55 = -(f[1] / p4(t2) + f[4] / t2);
  const double s56 = (f[4] / p2(t2) - f[2] / p4(t2));
  const double s57 = -(f[3] / p4(t2) + f[4] / p3(t2));
  const double s61 = -1.0, s62 = f[8], s63 = f[8], s64 = -1.0;
  const double s65 = (f[5] + f[8]), s66 = (f[8] - f[6]), s67 = (f[7] + f[8]);
  const double s71 = -1.0, s72 = (1.0 + f[4]), s73 = f[4];
  const double s74 = (f[4] - f[1]), s75 = (f[4] - f[2]), s76 = (f[4] - f[3]);
  const double q4 = std::pow(t2, 4.0), q2 = std::pow(t2, 2.0), q3 = std::pow(t2, 3.0);
  const double s81 = -1.0 / q4, s82 = (f[8] + (1.0 / q4)), s83 = f[8];
  const double s84 = (f[5] / q4 - f[8] / t2);
  const double s85 = (f[8] / q2 - f[6] / q4);
  const double s86 = (f[7] / q4 - f[8] / q3);
  const double s91 = -s66 * s51, s92 = s56 * s61, s93 = -(s54 * s66 + s56 * s62);
  const double s94 = (s56 * s64 + s66 * s52), s95 = (s56 * s63 - s66 * s53);
  const double s96 = (s56 * s65 - s66 * s55), s97 = (s56 * s67 - s66 * s57);
  const double sa1 = -s85 * s71, sa2 = -s85 * s72, sa3 = s75 * s81, sa4 = s75 * s82;
  const double sa5 = (s75 * s83 - s85 * s73), sa6 = (s75 * s84 - s85 * s74);
  const double sa7 = (s75 * s86 - s85 * s76);
  ca[0] = s97 * sa1 - sa7 * s91;
  ca[1] = s97 * sa2 + sa7 * s93;
  ca[2] = s95 * sa7 - sa5 * s97;
  ca[3] = s94 * sa7 + sa4 * s97;
  ca[4] = s97 * sa3 - sa7 * s92;
  ca[5] = s97 * sa6 - sa7 * s96;
  const double k1 = (1.0 + 2.0 * a) * (c + p2(c)) * t2;
  const double k2 = (1.0 + a) * (2.0 * c + p2(c));
  const double k3 = (1.0 + a);
  const double k4 = (1.0 + a) * p2(1.0 + c);
  const double k5 = p2(1.0 + a) * (c + p2(c)) * t2;
  const double k6 = p2(a) * t2 * (c + p2(c));
  const double k7 = (1.0 + a) * (c + p2(c)) * t2;
  ck[0] = k3;
  ck[1] = -k4;
  ck[2] = k1 + k2;
  ck[3] = -k5;
  ck[4] = k6;
  ck[5] = k7;
}

// same-direction tables use cell-width ratios, cross-direction tables the
// ratios of centre-to-centre spacings (NavierStokes.f90:42-110, :112-248)
static void upwind_tables(const std::vector<double> &d, int64_t n, bool cross,
                          std::vector<double> &ca, std::vector<double> &ck) {
  ca.assign((size_t)(6 * (n + 2)), 0.0);
  ck.assign((size_t)(6 * (n + 2)), 0.0);
  for (int64_t i = 2; i <= n; ++i) {
    double th[3];
    if (!cross) {
      th[0] = d[i - 1] / d[i - 2];
      th[1] = d[i] / d[i - 1];
      th[2] = d[i + 1] / d[i];
    } else {
      const double h1 = (i == 2) ? d[i - 2] : 0.5 * (d[i - 2] + d[i - 3]);
      const double h2 = 0.5 * (d[i - 1] + d[i - 2]);
      const double h3 = 0.5 * (d[i - 1] + d[i]);
      const double h4 = 0.5 * (d[i] + d[i + 1]);
      th[0] = h2 / h1;
      th[1] = h3 / h2;
      th[2] = h4 / h3;
    }
    upwind_column(th, &ca[(size_t)(6 * (i - 1))], &ck[(size_t)(6 * (i - 1))]);
  }
}

}  // namespace bcfd

using namespace bcfd;

extern "C" {

const char *biocfd_last_error(void) { return g_err; }

int biocfd_inject_alloc_failure(int nth) {
  BCFD_API_LOCK();
  if (nth < 0) { set_error("inject_alloc_failure: %d < 0", nth); return -1; }
  g_alloc_fail_in = nth;
  return 0;
}

int biocfd_init(int device) {
  BCFD_API_LOCK();
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_rt.init) return 0;
  int n = 0;
  BCFD_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    set_error("device %d not available (%d devices)", device, n);
    return -1;
  }
  BCFD_HIP(hipSetDevice(device));
  if (stress_scratch_init(device)) return -1;
  BCFD_HIP(hipStreamCreateWithFlags(&g_rt.stream, hipStreamNonBlocking));
  BCFD_HIP(hipHostMalloc((void **)&g_rt.h_pinned, 64 * sizeof(double), hipHostMallocDefault));
  BCFD_HIP(hipHostMalloc((void **)&g_rt.h_pinned_i, 64 * sizeof(int64_t), hipHostMallocDefault));
  g_rt.device = device;
  g_rt.init = true;
  return 0;
}

int biocfd_copy_probe(int64_t bytes, int reps) {
  BCFD_API_LOCK();
  if (ensure_init()) return -1;
  if (bytes <= 0 || reps <= 0) { set_error("copy_probe: bad size"); return -1; }
  void *a = nullptr, *b = nullptr;
  BCFD_HIP(hipMalloc(&a, (size_t)bytes));
  BCFD_HIP(hipMalloc(&b, (size_t)bytes));
  BCFD_HIP(hipMemsetAsync(a, 0, (size_t)bytes, g_rt.stream));
  for (int r = 0; r < reps; ++r)
    BCFD_HIP(hipMemcpyAsync(b, a, (size_t)bytes, hipMemcpyDeviceToDevice, g_rt.stream));
  BCFD_HIP(hipStreamSynchronize(g_rt.stream));
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

int biocfd_synchronize(void) {
  BCFD_API_LOCK();
  if (ensure_init()) return -1;
  BCFD_HIP(hipStreamSynchronize(g_rt.stream));
  return 0;
}

static void free_block(Block &b) {
  ad_release(b);
  sor_graphs_release(b);
  batch_graphs_release();
  for (auto &p : b.g) if (p) { (void)hipFree(p); p = nullptr; }
  for (auto &p : b.f) if (p) { (void)hipFree(p); p = nullptr; }
  for (auto &p : b.tag) if (p) { (void)hipFree(p); p = nullptr; }
  for (auto &p : b.A) if (p) { (void)hipFree(p); p = nullptr; }
  for (int t = 0; t < BTB_N; ++t) {
    if (b.ca[t]) (void)hipFree(b.ca[t]);
    if (b.ck[t]) (void)hipFree(b.ck[t]);
    if (b.ca32[t]) (void)hipFree(b.ca32[t]);
    if (b.ck32[t]) (void)hipFree(b.ck32[t]);
    b.ca[t] = b.ck[t] = nullptr;
    b.ca32[t] = b.ck32[t] = nullptr;
    if (b.rr[t]) (void)hipFree(b.rr[t]);
    b.rr[t] = nullptr;
  }
  for (auto &p : b.h32) if (p) { (void)hipFree(p); p = nullptr; }
  if (b.mom_fix) { (void)hipFree(b.mom_fix); b.mom_fix = nullptr; b.cap_mom_fix = 0; }
  if (b.sor) (void)hipFree(b.sor);
  if (b.scan_tmp) (void)hipFree(b.scan_tmp);
  if (b.list_tmp) (void)hipFree(b.list_tmp);
  for (int c = 0; c < 2; ++c) {
    if (b.pcs[c]) (void)hipFree(b.pcs[c]);
    if (b.bs[c]) (void)hipFree(b.bs[c]);
  }
  if (b.pc2) (void)hipFree(b.pc2);
  for (int q = 0; q < 2; ++q) {
    if (b.e32[q]) (void)hipFree(b.e32[q]);
    if (b.r32[q]) (void)hipFree(b.r32[q]);
  }
  if (b.partials) (void)hipFree(b.partials);
  for (auto &p : b.mv.dum) if (p) (void)hipFree(p);
  if (b.out_stage) (void)hipFree(b.out_stage);
  IbData &ib = b.ib;
  for (auto &p : ib.el) if (p) (void)hipFree(p);
  if (ib.surf) (void)hipFree(ib.surf);
  for (auto &p : ib.node) if (p) (void)hipFree(p);
  for (auto &p : ib.elem) if (p) (void)hipFree(p);
  for (void *p : {ib.bin_grid, (void *)ib.bin_start, (void *)ib.bin_elem, (void *)ib.bin_fill})
    if (p) (void)hipFree(p);
  if (ib.ib_ijk) (void)hipFree(ib.ib_ijk);
  if (ib.retag_bits) (void)hipFree(ib.retag_bits);
  for (auto &p : ib.nel) if (p) (void)hipFree(p);
  for (auto &p : ib.normdis) if (p) (void)hipFree(p);
  if (ib.ib_map) (void)hipFree(ib.ib_map);
  if (ib.ts_ijk) (void)hipFree(ib.ts_ijk);
  if (ib.index_ts) (void)hipFree(ib.index_ts);
  for (auto &p : ib.ghost) if (p) (void)hipFree(p);
  if (ib.solid_ijk) (void)hipFree(ib.solid_ijk);
  if (ib.task_out) (void)hipFree(ib.task_out);
  if (ib.task_done) (void)hipFree(ib.task_done);
  if (ib.stress_work) (void)hipFree(ib.stress_work);
  for (void *p : {(void *)ib.ts_mask, (void *)ib.ts_list, (void *)ib.ts_best, (void *)ib.ts_qmap, (void *)ib.ts_adjm,
                  (void *)ib.ts_src, (void *)ib.ts_val})
    if (p) (void)hipFree(p);
  b = Block();
}

int biocfd_finalize(void) {
  BCFD_API_LOCK();
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_rt.init) return 0;
  (void)hipStreamSynchronize(g_rt.stream);
  for (int i = 1; i <= kMaxBlocks; ++i)
    if (g_rt.blk[i].live) free_block(g_rt.blk[i]);
  for (int i = 1; i <= kMaxInterfaces; ++i) {
    Interface &it = g_rt.itf[i];
    for (auto &row : it.det)
      for (auto &p : row)
        if (p) (void)hipFree(p);
    it = Interface();
  }
  (void)hipHostFree(g_rt.h_pinned);
  (void)hipHostFree(g_rt.h_pinned_i);
  (void)hipStreamDestroy(g_rt.stream);
  g_rt = Runtime();
  return 0;
}

}  // extern "C"

namespace bcfd {
// the device arrays of a new block, cleared (ib_map: -1)
static int block_alloc(Block &b) {
  for (int q = 0; q < BF_NFIELD; ++q) {
    if (q == BF_PCO) continue;
    if (dev_alloc(&b.f[q], b.n2)) return -1;
    BCFD_HIP(hipMemsetAsync(b.f[q], 0, sizeof(double) * b.n2, g_rt.stream));
  }
  for (int t = 0; t < BT_NTAG; ++t) {
    int64_t n = tag_size(b, t);
    if (dev_alloc(&b.tag[t], n)) return -1;
    BCFD_HIP(hipMemsetAsync(b.tag[t], 0, n, g_rt.stream));
  }
  if (dev_alloc(&b.ib.ib_map, b.n2)) return -1;
  BCFD_HIP(hipMemsetAsync(b.ib.ib_map, 0xff, sizeof(int) * b.n2, g_rt.stream));
  if (dev_alloc(&b.sor, 1)) return -1;
  BCFD_HIP(hipMemsetAsync(b.sor, 0, sizeof(SorState), g_rt.stream));
  return 0;
}
}  // namespace bcfd

extern "C" {

int biocfd_block_create(int id, int64_t nx, int64_t ny, int64_t nz, double omega, int64_t cintp,
                        double dx, double dy, double dz) {
  BCFD_API_LOCK();
  if (ensure_init()) return -1;
  if (id < 1 || id > kMaxBlocks) {
    set_error("block id %d out of range 1..%d", id, kMaxBlocks);
    return -1;
  }
  if (nx < 3 || ny < 3 || nz < 3 || nx > 4096 || ny > 4096 || nz > 4096) {
    set_error("block %d: unsupported size %lldx%lldx%lld", id, (long long)nx, (long long)ny,
              (long long)nz);
    return -1;
  }
  Block &b = g_rt.blk[id];
  if (b.live) free_block(b);
  b.d.nx = (int)nx; b.d.ny = (int)ny; b.d.nz = (int)nz;
  b.d.NX = (int)nx + 2; b.d.NY = (int)ny + 2; b.d.NZ = (int)nz + 2;
  b.d.TX = (int)nx + 3; b.d.TY = (int)ny + 3;
  b.n2 = (int64_t)b.d.NX * b.d.NY * b.d.NZ;
  b.n3 = (int64_t)(nx + 3) * (ny + 3) * (nz + 3);
  b.omega = omega; b.cintp = cintp; b.dx = dx; b.dy = dy; b.dz = dz;
  if (block_alloc(b)) {
    // a failed create leaves no block and no device memory behind
    (void)hipStreamSynchronize(g_rt.stream);
    free_block(b);
    return -1;
  }
  b.live = true;
  return 0;
}

int biocfd_block_destroy(int id) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  (void)hipStreamSynchronize(g_rt.stream);
  free_block(*b);
  return 0;
}

static int64_t grid_len(const Block &b, int which) {
  const int64_t n[3] = {b.d.nx, b.d.ny, b.d.nz};
  switch (which) {
    case BG_X1: case BG_XU: return n[0] + 3;
    case BG_Y1: case BG_YV: return n[1] + 3;
    case BG_Z1: case BG_ZW: return n[2] + 3;
    case BG_DX: case BG_XV: case BG_XW: case BG_XP: return n[0] + 2;
    case BG_DY: case BG_YU: case BG_YW: case BG_YP: return n[1] + 2;
    default: return n[2] + 2;
  }
}

int biocfd_grid_upload(int id, int which, const double *host, int64_t n) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (which < 0 || which >= BG_NGRID || n != grid_len(*b, which)) {
    set_error("grid %d: expected %lld entries, got %lld", which,
              (long long)(which >= 0 && which < BG_NGRID ? grid_len(*b, which) : -1), (long long)n);
    return -1;
  }
  b->hg[which].assign(host, host + n);
  if (dev_alloc(&b->g[which], n)) return -1;
  BCFD_HIP(hipMemcpyAsync(b->g[which], host, sizeof(double) * n, hipMemcpyHostToDevice, g_rt.stream));
  BCFD_HIP(hipStreamSynchronize(g_rt.stream));
  return 0;
}

int biocfd_field_upload(int id, int field, const double *host) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (field < 0 || field >= BF_NFIELD) { set_error("bad field %d", field); return -1; }
  if (field == BF_PCO) field = BF_PC;
  BCFD_HIP(hipMemcpyAsync(b->f[field], host, sizeof(double) * b->n2, hipMemcpyHostToDevice, g_rt.stream));
  BCFD_HIP(hipStreamSynchronize(g_rt.stream));
  return 0;
}

int biocfd_field_download(int id, int field, double *host) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (field < 0 || field >= BF_NFIELD) { set_error("bad field %d", field); return -1; }
  // pco equals pc after every reference-visible point (PcorVcor.F90:363-372,
  // CoarseInterp.f90:413-414, FineInterp_bound.f90:90-91)
  if (field == BF_PCO) field = BF_PC;
  BCFD_HIP(hipMemcpyAsync(host, b->f[field], sizeof(double) * b->n2, hipMemcpyDeviceToHost, g_rt.stream));
  BCFD_HIP(hipStreamSynchronize(g_rt.stream));
  return 0;
}

int biocfd_tag_upload(int id, int tag, const int64_t *host) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (tag < 0 || tag >= BT_NTAG) { set_error("bad tag %d", tag); return -1; }
  int64_t n = tag_size(*b, tag);
  std::vector<uint8_t> t((size_t)n);
  for (int64_t q = 0; q < n; ++q) {
    if (host[q] < 0 || host[q] > 255) { set_error("tag value %lld out of range", (long long)host[q]); return -1; }
    t[q] = (uint8_t)host[q];
  }
  BCFD_HIP(hipMemcpyAsync(b->tag[tag], t.data(), n, hipMemcpyHostToDevice, g_rt.stream));
  BCFD_HIP(hipStreamSynchronize(g_rt.stream));
  return 0;
}

int biocfd_tag_download(int id, int tag, int64_t *host) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (tag < 0 || tag >= BT_NTAG) { set_error("bad tag %d", tag); return -1; }
  int64_t n = tag_size(*b, tag);
  std::vector<uint8_t> t((size_t)n);
  BCFD_HIP(hipMemcpyAsync(t.data(), b->tag[tag], n, hipMemcpyDeviceToHost, g_rt.stream));
  BCFD_HIP(hipStreamSynchronize(g_rt.stream));
  for (int64_t q = 0; q < n; ++q) host[q] = t[q];
  return 0;
}

int biocfd_counts(int id, int64_t out[BC_N]) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  for (int q = 0; q < BC_N; ++q) out[q] = b->counts[q];
  return 0;
}

int biocfd_sor_state(int id, int64_t *n_iter, double *derr2, double *derr_std_st) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  SorState s;
  BCFD_HIP(hipMemcpyAsync(&s, b->sor, sizeof(s), hipMemcpyDeviceToHost, g_rt.stream));
  BCFD_HIP(hipStreamSynchronize(g_rt.stream));
  if (n_iter) *n_iter = s.n_iter;
  if (derr2) *derr2 = s.derr2;
  if (derr_std_st) *derr_std_st = s.derr_std_st;
  return 0;
}

int biocfd_coefficient_matrix(int id, int is_coarse) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (b->hg[BG_XP].empty() || b->hg[BG_YP].empty() || b->hg[BG_ZP].empty()) {
    set_error("coefficient_matrix: cell-centre grids not uploaded");
    return -1;
  }
  const int64_t n[3] = {b->d.nx, b->d.ny, b->d.nz};
  const int gi[3] = {BG_XP, BG_YP, BG_ZP};
  for (int a = 0; a < 3; ++a) {
    laplace_rows(b->hg[gi[a]], n[a], b->hA[a]);
    if (is_coarse) fold(b->hA[a], n[a], a == 0);  // x: inlet/outlet Dirichlet, y/z: walls Neumann
    if (dev_alloc(&b->A[a], 3 * n[a])) return -1;
    BCFD_HIP(hipMemcpyAsync(b->A[a], b->hA[a].data(), sizeof(double) * 3 * n[a],
                            hipMemcpyHostToDevice, g_rt.stream));
  }
  BCFD_HIP(hipStreamSynchronize(g_rt.stream));
  return poisson_block_setup(*b);
}

int biocfd_non_uni_coeff(int id) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (b->hg[BG_DX].empty() || b->hg[BG_DY].empty() || b->hg[BG_DZ].empty()) {
    set_error("non_uni_coeff: cell widths not uploaded");
    return -1;
  }
  const std::vector<double> *dd[3] = {&b->hg[BG_DX], &b->hg[BG_DY], &b->hg[BG_DZ]};
  const int64_t n[3] = {b->d.nx, b->d.ny, b->d.nz};
  // table -> (axis, cross?) ; uw == uv, vw == vu, wv == wu
  const int axis[BTB_N] = {0, 1, 2, 0, 0, 1, 1, 2, 2};
  const bool cross[BTB_N] = {false, false, false, true, true, true, true, true, true};
  for (int t = 0; t < BTB_N; ++t) {
    upwind_tables(*dd[axis[t]], n[axis[t]], cross[t], b->hca[t], b->hck[t]);
    const int64_t m = 6 * (n[axis[t]] + 2);
    if (dev_alloc(&b->ca[t], m) || dev_alloc(&b->ck[t], m)) return -1;
    BCFD_HIP(hipMemcpyAsync(b->ca[t], b->hca[t].data(), sizeof(double) * m, hipMemcpyHostToDevice, g_rt.stream));
    BCFD_HIP(hipMemcpyAsync(b->ck[t], b->hck[t].data(), sizeof(double) * m, hipMemcpyHostToDevice, g_rt.stream));
  }
  // the upwind denominators of each table and their reciprocals (the
  // momentum kernel divides by them through a correction of the product
  // with the reciprocal; NavierStokes.f90:619-647's ca6 hA and 2 ck6 hB with
  // the stencil's spacings: a same-direction table at cell index i takes
  // hA = d(i), hB = d(i-1), a cross table the centre spacings
  // 0.5 (d(i-1) + d(i)) and 0.5 (d(i-1) + d(i-2)))
  for (int t = 0; t < BTB_N; ++t) {
    const std::vector<double> &d = *dd[axis[t]];
    const int64_t nn = n[axis[t]];
    std::vector<double> rr((size_t)(4 * (nn + 2)), 0.0);
    bool ok = true;
    for (int64_t i = 2; i <= nn; ++i) {
      const double hA = cross[t] ? 0.5 * (d[i - 1] + d[i]) : d[i];
      const double hB = cross[t] ? 0.5 * (d[i - 1] + d[i - 2]) : d[i - 1];
      const double b1 = b->hca[t][(size_t)(6 * (i - 1) + 5)] * hA;
      const double b2 = 2.0 * b->hck[t][(size_t)(6 * (i - 1) + 5)] * hB;
      for (double x : {b1, b2})
        ok = ok && std::fabs(x) >= 0x1p-40 && std::fabs(x) <= 0x1p40;
      double *row = &rr[(size_t)(4 * (i - 1))];
      row[0] = b1;
      row[1] = 1.0 / b1;
      row[2] = b2;
      row[3] = 1.0 / b2;
    }
    if (b->rr[t]) { (void)hipFree(b->rr[t]); b->rr[t] = nullptr; }
    if (!ok) continue;
    if (dev_alloc(&b->rr[t], (int64_t)rr.size())) return -1;
    BCFD_HIP(hipMemcpy(b->rr[t], rr.data(), sizeof(double) * rr.size(), hipMemcpyHostToDevice));
  }
  // FP32 copies of the tables and spacings for the FP32 momentum tendency
  std::vector<float> f32;
  auto put32 = [&](float **dst, const std::vector<double> &src) -> int {
    f32.assign(src.begin(), src.end());
    if (dev_alloc(dst, (int64_t)f32.size())) return -1;
    BCFD_HIP(hipMemcpy(*dst, f32.data(), sizeof(float) * f32.size(), hipMemcpyHostToDevice));
    return 0;
  };
  for (int t = 0; t < BTB_N; ++t)
    if (put32(&b->ca32[t], b->hca[t]) || put32(&b->ck32[t], b->hck[t])) return -1;
  for (int a = 0; a < 3; ++a)
    if (put32(&b->h32[a], *dd[a])) return -1;
  BCFD_HIP(hipStreamSynchronize(g_rt.stream));
  return 0;
}

int biocfd_coef_download(int id, int axis, double *host, int64_t n) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (axis < 0 || axis > 2 || b->hA[axis].size() != (size_t)(3 * n)) {
    set_error("coef_download: axis %d size mismatch", axis);
    return -1;
  }
  memcpy(host, b->hA[axis].data(), sizeof(double) * 3 * n);
  return 0;
}

int biocfd_table_download(int id, int which, int is_ck, double *host, int64_t ncols) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (which < 0 || which >= BTB_N) { set_error("bad table %d", which); return -1; }
  const std::vector<double> &v = is_ck ? b->hck[which] : b->hca[which];
  if (v.size() != (size_t)(6 * ncols)) { set_error("table size mismatch"); return -1; }
  memcpy(host, v.data(), sizeof(double) * v.size());
  return 0;
}

int biocfd_body_create(int id, int64_t n_nodes, int64_t n_elems, const int64_t *p1,
                       const int64_t *p2, const int64_t *p3, const int64_t *surf,
                       const int64_t box[6]) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  IbData &ib = b->ib;
  if (n_nodes < 3 || n_elems < 1) { set_error("body: empty mesh"); return -1; }
  const int lim[3] = {b->d.nx + 1, b->d.ny + 1, b->d.nz + 1};
  for (int a = 0; a < 3; ++a) {
    if (box[2 * a] < 2 || box[2 * a + 1] > lim[a] || box[2 * a] > box[2 * a + 1]) {
      set_error("body: search box axis %d [%lld,%lld] outside 2..%d", a, (long long)box[2 * a],
                (long long)box[2 * a + 1], lim[a]);
      return -1;
    }
    ib.search[2 * a] = (int)box[2 * a];
    ib.search[2 * a + 1] = (int)box[2 * a + 1];
  }
  ib.n_nodes = n_nodes;
  ib.n_elems = n_elems;
  const int64_t *src[3] = {p1, p2, p3};
  std::vector<int> tmp((size_t)n_elems);
  for (int c = 0; c < 3; ++c) {
    for (int64_t e = 0; e < n_elems; ++e) {
      if (src[c][e] < 1 || src[c][e] > n_nodes) { set_error("body: element %lld has bad node", (long long)e + 1); return -1; }
      tmp[e] = (int)src[c][e];
    }
    if (dev_alloc(&ib.el[c], n_elems)) return -1;
    BCFD_HIP(hipMemcpy(ib.el[c], tmp.data(), sizeof(int) * n_elems, hipMemcpyHostToDevice));
  }
  ib.h_surf.resize((size_t)n_elems);
  for (int64_t e = 0; e < n_elems; ++e) ib.h_surf[e] = (int)surf[e];
  if (dev_alloc(&ib.surf, n_elems)) return -1;
  BCFD_HIP(hipMemcpy(ib.surf, ib.h_surf.data(), sizeof(int) * n_elems, hipMemcpyHostToDevice));
  for (int c = 0; c < 3; ++c)
    if (dev_alloc(&ib.node[c], n_nodes)) return -1;
  for (int q = 0; q < BE_NELEM; ++q)
    if (dev_alloc(&ib.elem[q], n_elems)) return -1;
  ib.present = true;
  return 0;
}

// ---- body kinematics (host arrays) ------------------------------------------
enum {
  S_XSHIFT = 0, S_YSHIFT, S_ZSHIFT, S_A0, S_YAMP, S_BFREQ, S_XPTH1, S_XPTH2, S_YPTH2, S_XCHG,
  S_YCHG, S_YT, S_YDOT, S_YDDOT, S_XT, S_XDOT, S_THETADOT, S_THETADDOT, S_THETADOT1,
  S_THETADDOT1, S_THETADOT2, S_THETADDOT2, S_ALPHADOT, S_PIVX, S_PIVY, S_PIVZ, S_XMOVE, S_YMOVE,
  S_ZMOVE, S_INITY, S_NXTY, S_INITX, S_NXTX, S_INITZ, S_NXTZ
};
static const double kPi = 3.141592653589793115997963468544185161590576171875;
static const double kDrift = 0.05;  // x drift per step in units of dxmin (Search.F90:12)

// flapping nodes (IDs 51/52) rotate about the x axis through piv_pt by +-aoa
static void place_nodes(int64_t nn, const double *xn, const double *yn, const double *zn,
                        const int64_t *id, double aoa1, double piv, double ox, double oy,
                        double oz, double *x1, double *y1, double *z1) {
  const double aoa2 = -aoa1;
  for (int64_t i = 0; i < nn; ++i) {
    double xr = xn[i], yr = yn[i], zr = zn[i];
    if (id[i] == 51 || id[i] == 52) {
      const double a = (id[i] == 51) ? aoa1 : aoa2;
      zr = zn[i] * cos(a) + yn[i] * sin(a) + piv - piv * cos(a);
      yr = -zn[i] * sin(a) + yn[i] * cos(a) + piv * sin(a);
    }
    x1[i] = xr + ox;
    y1[i] = yr + oy;
    z1[i] = zr + oz;
  }
}

int biocfd_shift_surface_nodes_initial(int64_t nn, const double *xn, const double *yn,
                                       const double *zn, const int64_t *id, double aoa,
                                       double piv_pt, int64_t ita, double dt, double dxmin,
                                       double totime, double *s, double *x1, double *y1,
                                       double *z1) {
  BCFD_API_LOCK();
  const double aoa1 = aoa * kPi / 180.0;
  const double bdfr = 15.0, bdy = 15.0 * dxmin;
  s[S_XMOVE] = s[S_YMOVE] = s[S_ZMOVE] = 0.0;
  s[S_A0] = s[S_A0] * kPi / 180.0;
  const double angt = 2.0 * kPi * s[S_BFREQ];
  s[S_XPTH1] = s[S_XSHIFT] - ((double)ita * dxmin * kDrift);
  s[S_XPTH2] = s[S_XSHIFT] - ((double)ita * dxmin * kDrift);
  s[S_YPTH2] = (s[S_YAMP]) * sin(angt * s[S_XSHIFT]);
  s[S_PIVX] = s[S_XSHIFT]; s[S_PIVY] = s[S_YSHIFT]; s[S_PIVZ] = s[S_ZSHIFT];
  s[S_THETADOT] = s[S_THETADDOT] = s[S_THETADOT1] = s[S_THETADOT2] = 0.0;
  s[S_ALPHADOT] = s[S_THETADDOT1] = s[S_THETADDOT2] = 0.0;
  s[S_YT] = bdy * sin(2.0 * kPi * bdfr * totime);
  s[S_YDOT] = angt * bdy * cos(2.0 * kPi * bdfr * totime);
  s[S_YDDOT] = -angt * angt * bdy * sin(2.0 * kPi * bdfr * totime);
  s[S_XT] = s[S_XSHIFT] - ((double)ita * dxmin * kDrift);
  s[S_XDOT] = -(dxmin * kDrift) / dt;
  s[S_INITY] = s[S_NXTY] = s[S_YSHIFT];
  s[S_INITX] = s[S_NXTX] = s[S_XSHIFT];
  s[S_INITZ] = s[S_NXTZ] = s[S_ZSHIFT];
  place_nodes(nn, xn, yn, zn, id, aoa1, piv_pt, s[S_XSHIFT], s[S_YSHIFT], s[S_ZSHIFT], x1, y1, z1);
  return 0;
}

int biocfd_compute_surface_variables(int64_t nn, const double *xn, const double *yn,
                                     const double *zn, const int64_t *id, double phase,
                                     double piv_pt, int64_t ita, double dt, double dxmin,
                                     double freq, double totime, double *s, double *x1,
                                     double *y1, double *z1) {
  BCFD_API_LOCK();
  const double ph = 2.0 * kPi * freq * (totime + dt) + phase;
  const double aoa1 = (s[S_A0]) * sin(ph);
  const double wtheta = 2.0 * kPi * freq;
  const double bdy = s[S_YAMP];
  const double angt = 2.0 * kPi * s[S_BFREQ];
  s[S_XPTH2] = s[S_XSHIFT] - ((double)ita * dxmin * kDrift);
  s[S_YPTH2] = (s[S_YAMP]) * sin(angt * s[S_XPTH2]);
  s[S_XCHG] = s[S_XPTH2] - s[S_XPTH1];
  s[S_XPTH1] = s[S_XPTH2];
  s[S_THETADOT1] = wtheta * s[S_A0] * cos(ph);
  s[S_THETADDOT1] = -wtheta * wtheta * s[S_A0] * sin(ph);
  s[S_THETADOT2] = -s[S_THETADOT1];
  s[S_THETADDOT2] = -s[S_THETADDOT1];
  s[S_YT] = bdy * sin(angt * (totime));
  s[S_YDOT] = angt * bdy * cos(angt * (totime));
  s[S_YDDOT] = -angt * angt * bdy * sin(angt * (totime));
  s[S_XT] = s[S_XPTH2];
  s[S_XDOT] = -(dxmin * kDrift) / dt;
  s[S_YCHG] = s[S_YT] - bdy * sin(angt * (totime - dt));
  s[S_YMOVE] = s[S_YT];
  s[S_XMOVE] = s[S_XCHG];
  s[S_ZMOVE] = 0.0;
  s[S_PIVY] = s[S_PIVY] + s[S_YCHG];
  s[S_PIVX] = s[S_PIVX] + s[S_XMOVE];
  s[S_PIVZ] = s[S_PIVZ] + s[S_ZMOVE];
  s[S_NXTY] = s[S_NXTY] + s[S_YCHG];
  s[S_NXTX] = s[S_NXTX] + s[S_XMOVE];
  // x1 = xr + xpth2 ; y1 = (yr + yshift) + ymove ; z1 = (zr + zshift) + zmove
  place_nodes(nn, xn, yn, zn, id, aoa1, piv_pt, s[S_XPTH2], s[S_YSHIFT], s[S_ZSHIFT], x1, y1, z1);
  for (int64_t i = 0; i < nn; ++i) {
    y1[i] = y1[i] + s[S_YMOVE];
    z1[i] = z1[i] + s[S_ZMOVE];
  }
  return 0;
}

}  // extern "C"
