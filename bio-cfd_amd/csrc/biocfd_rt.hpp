// Internal header of the gfx950 runtime behind include/biocfd_hip.h.
//
// Device layout: every field stays in the reference's Fortran layout
// (column-major, 1-based, one ghost layer) so host<->device transfers are a
// single copy and the i index is the contiguous (coalesced) one.  Cell tags
// live on the device as uint8 (the reference keeps int64); they are widened
// only when downloaded.  Index lists are int32 triplets, built by ordered
// (stable) scans so their order is the reference's serial i-fastest order.
#pragma once

#include <mutex>

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "biocfd_hip.h"

namespace bcfd {

constexpr int kMaxBlocks = 64;
constexpr int kMaxInterfaces = 64;

// Every C entry point holds this lock while it runs: host threads may call
// the library concurrently (the reference drives its blocks from OpenMP
// threads), and their work is queued in call order on the library's one
// in-order stream.  Recursive, as some entry points call others.
std::recursive_mutex &api_mutex();
#define BCFD_API_LOCK() std::lock_guard<std::recursive_mutex> bcfd_api_lock_(::bcfd::api_mutex())

void set_error(const char *fmt, ...);

#define BCFD_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      ::bcfd::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return -1;                                                                         \
    }                                                                                    \
  } while (0)

#define BCFD_CHECK_LAUNCH()                                                              \
  do {                                                                                   \
    hipError_t e_ = hipGetLastError();                                                   \
    if (e_ != hipSuccess) {                                                              \
      ::bcfd::set_error("%s:%d launch: %s", __FILE__, __LINE__, hipGetErrorString(e_));  \
      return -1;                                                                         \
    }                                                                                    \
  } while (0)

// Per-block scalars a kernel needs, passed by value.
struct Dims {
  int nx, ny, nz;     // interior cells
  int NX, NY, NZ;     // nx+2 ...
  int TX, TY;         // nx+3, ny+3 (tall tag arrays)
};

__host__ __device__ inline int64_t ix3(const Dims &d, int i, int j, int k) {
  return (int64_t)(i - 1) + (int64_t)d.NX * ((int64_t)(j - 1) + (int64_t)d.NY * (int64_t)(k - 1));
}
__host__ __device__ inline int64_t ixt(const Dims &d, int i, int j, int k) {
  return (int64_t)(i - 1) + (int64_t)d.TX * ((int64_t)(j - 1) + (int64_t)d.TY * (int64_t)(k - 1));
}

// SOR control state in device memory (one per block).
struct SorState {
  long long n_iter;              // nIterPcor (accumulates within one poissonSolver call)
  double derr4;                  // last update norm (11111 at call start)
  double derr2;                  // block%derr2
  unsigned long long max_bits;   // running max of |pc-pco| (bit pattern of a double >= 0)
  int done;
  int passes;                    // fused launches of the current SOR call
  double derr_std_st;            // steady-state norm
  unsigned long long std_bits;
  long long call_iters;          // iterations executed by the current SOR call
  int exhausted;                 // SOR calls since the last nIterPcor reset that ended on their
                                 // iteration budget, not their tolerance
  int rollback;                  // the call stopped on the first iteration of a two-iteration pass:
                                 // that iteration is redone alone from the pass's source
};

struct IbData {
  int64_t n_nodes = 0, n_elems = 0;
  int search[6] = {0, 0, 0, 0, 0, 0};
  int *el[3] = {nullptr, nullptr, nullptr};   // 1-based node numbers
  int *surf = nullptr;                        // per element
  double *node[3] = {nullptr, nullptr, nullptr};  // xnode1..znode1
  double *elem[BE_NELEM] = {};                // xcent..element_length
  std::vector<int> h_surf;
  // uniform grid of the element centroids, counting-sorted by bin (rebuilt
  // whenever the centroids are): bin grid, first entry of each bin, element
  // of each entry, per-bin counters
  void *bin_grid = nullptr;
  int *bin_start = nullptr, *bin_elem = nullptr, *bin_fill = nullptr;
  bool bins_ready = false;
  // one bit per cell of the block: the neighbour cells selective retagging
  // has claimed in the current call (all clear between calls)
  unsigned *retag_bits = nullptr;
  int64_t retag_words = 0;
  // per IB cell (capacity grows)
  int64_t cap_ib = 0;
  int64_t n_mapped = 0;                       // entries of ib_ijk currently in ib_map
  int *ib_ijk = nullptr;                      // 3*cap: i list, j list, k list
  int *nel[BN_NNEL] = {};
  double *normdis[BN_NNEL] = {};
  int *ib_map = nullptr;                      // n2: IB ordinal (0-based) or -1
  // TS cells
  int64_t cap_ts = 0;
  int *ts_ijk = nullptr;
  int *index_ts = nullptr;                    // 1-based IB ordinal
  double *ghost[BGH_N] = {};
  // solid list
  int64_t cap_solid = 0;
  int *solid_ijk = nullptr;
  // nsMomentum2order's TS pass: touched and deferred masks (2 n2 bytes), the
  // deferred cells and the ghost each of their targets takes
  uint8_t *ts_mask = nullptr;
  int64_t cap_ts_mask = 0;
  int *ts_list = nullptr, *ts_best = nullptr;
  int64_t cap_ts_list = 0, cap_ts_best = 0;
  // the parallel form of that pass: list index of each deferred cell (n2,
  // -1 elsewhere between calls) and the TS-adjacency mask of each entry
  int *ts_qmap = nullptr;
  int64_t cap_ts_qmap = 0;
  uint8_t *ts_adjm = nullptr;
  int64_t cap_ts_adjm = 0;
  // the deferred cells' 49 operands resolved: source words (where each is
  // read) and two tables of values (operands, then tangents or adjoints)
  int64_t *ts_src = nullptr;
  double *ts_val = nullptr;
  int64_t cap_ts_src = 0, cap_ts_val = 0;
  // stress_calculation: per-element terms, the coefficients and the miss counter
  double *stress_work = nullptr;
  int64_t cap_stress = 0;
  // forcing scratch (per task)
  int64_t cap_task = 0;
  double *task_out = nullptr;
  int *task_done = nullptr;
  int64_t cap_done = 0;
  bool present = false;
};

// The last block move of a fine block (Blocks.f90:64-71): the shift in
// cells per axis, the index ranges of the kept overlap before the move (cpy)
// and after it (cpy_mv), x start/end, y start/end, z start/end, and the
// fields u, v, w, p kept for the shifted copy (device, allocated at the
// first move).
struct MoveState {
  int64_t amt[3] = {0, 0, 0};
  int64_t cpy[6] = {0, 0, 0, 0, 0, 0};
  int64_t cpy_mv[6] = {0, 0, 0, 0, 0, 0};
  double *dum[4] = {nullptr, nullptr, nullptr, nullptr};
};

// Sensitivities of the time step for the tangent-linear and adjoint passes
// (FP64, block layout): s[q] for the fields BF_U .. BF_P (u, v, w, ut, vt,
// wt, p), the ghost arrays', two scratch arrays of the transposed projection, the per-task
// adjoints of the forcing routines and a kinematic accumulator (BK_N).
// What one forcing call read, kept while recording: the arrays it
// interpolates at the call's start (p or ut, vt, wt) and, for the in-place
// routines, its task outputs.
struct ForceRec {
  double *arr[4] = {};
  double *tout = nullptr;
  int64_t cap_tout = 0, ntask = 0;
  bool live = false;
};

struct AdWork {
  double *s[7] = {};
  double *gh[BGH_N] = {};   // the thin-surface ghost arrays' sensitivities (cap_gh entries each)
  int64_t cap_gh = 0;
  double *scr[3] = {};
  double *ybar = nullptr;
  int64_t cap_ybar = 0;
  double *kin = nullptr;
  double *addend = nullptr;  // per-task geometric tangents of a forcing call
  int64_t cap_addend = 0;
  // geometric sensitivities per element, 7 x n_elems each (centroid x, y, z,
  // normal x, y, z, element length): adjoints (gbar) and tangents (gdot);
  // node adjoints, 3 x n_nodes
  double *gbar = nullptr, *gdot = nullptr, *nbar = nullptr;
  bool gdot_set = false;
  // forcing-call records (slot * 6 + routine) and the slot being recorded
  std::vector<ForceRec> recs;
  int rec_slot = -1;
  // the transposed SOR's launches as one graph (fixed count: no host polls)
  hipGraphExec_t sor_exec = nullptr;
  long long sor_key[3] = {0, 0, 0};   // iterations, dt bits, cell tags
};

// a captured chunk of SOR passes of one block (k_pcor.hip, graph_chunk)
struct SorGraph {
  static constexpr int kKey = 10, kPtrs = 9;
  long long key[kKey];
  const void *ptrs[kPtrs];       // the block's buffers the launches name
  long long d_iters, d_q;        // iterations and passes the chunk runs
  long long end_it, end_q;       // its last end kernel, relative to the chunk's start
  hipGraphExec_t exec = nullptr;
};

struct Block {
  bool live = false;
  Dims d{};
  int64_t n2 = 0, n3 = 0;
  double omega = 1.0, dx = 0, dy = 0, dz = 0;
  int64_t cintp = 3;
  std::vector<double> hg[BG_NGRID];   // host copies of the grid arrays
  double *g[BG_NGRID] = {};           // device grid arrays
  double *f[BF_NFIELD] = {};          // device fields (pco is not stored: it equals pc)
  uint8_t *tag[BT_NTAG] = {};
  double *A[3] = {};                  // Acx, Acy, Acz (n,3)
  std::vector<double> hA[3];
  double *ca[BTB_N] = {}, *ck[BTB_N] = {};
  float *ca32[BTB_N] = {}, *ck32[BTB_N] = {}, *h32[3] = {};   // FP32 copies (FP32 momentum tendency)
  // per table and cell index: the upwind denominators b1 = ca6 hA, b2 = (2 ck6) hB
  // and their reciprocals rounded once {b1, 1/b1, b2, 1/b2} (k_momentum.hip);
  // null where a denominator is outside [2^-40, 2^40]
  double *rr[BTB_N] = {};
  int *mom_fix = nullptr;             // cells the momentum march evaluates again (count, then cells)
  int64_t cap_mom_fix = 0;
  std::vector<double> hca[BTB_N], hck[BTB_N];
  SorState *sor = nullptr;            // device
  int64_t counts[BC_N] = {0, 0, 0, 0, 0, 0};
  IbData ib;
  // scan scratch
  int64_t *scan_tmp = nullptr;
  int64_t scan_cap = 0;
  int *list_tmp = nullptr;
  int64_t list_cap = 0;
  // colour-split pc and b used inside the SOR iterations
  double *pcs[2] = {nullptr, nullptr}, *bs[2] = {nullptr, nullptr};
  int64_t split_n = 0;
  double *pc2 = nullptr;              // second pc array of the fused SOR iterations
  int64_t pc2_n = 0;
  int sor_variant = -1, sor_planes = 0;   // this block's own FP64 SOR kernel (-1: the library's)
  // mixed-precision solve: FP32 correction pair (full layout) and FP32
  // residual halves (colour-split)
  float *e32[2] = {nullptr, nullptr}, *r32[2] = {nullptr, nullptr};
  int64_t mx_n = 0;
  std::vector<SorGraph> sor_graphs;   // captured SOR pass chunks
  double *partials = nullptr;         // per-workgroup maxima of the reductions
  int64_t cap_partials = 0;
  MoveState mv;
  // output view staging (cell-centre velocities, p, tags of the interior)
  double *out_stage = nullptr;
  int64_t out_cap = 0;
  AdWork ad;
};

struct Interface {
  bool live = false;
  int a = 0, b = 0;                    // 1-based block numbers
  int64_t a_msh = 1, b_msh = 1;
  double ext[6] = {0, 0, 0, 0, 0, 0};
  double ext_new[6] = {0, 0, 0, 0, 0, 0};  // the overlap a block move keeps
  // det tables [comp p,u,v,w][axis x,y,z]: host (3,cap) and device copies
  std::vector<int64_t> hdet[4][3];
  int64_t counter[4][3] = {};
  int *det[4][3] = {};                 // device int32 (3, counter)
};

struct Runtime {
  bool init = false;
  int device = 0;
  hipStream_t stream = nullptr;
  Block blk[kMaxBlocks + 1];
  Interface itf[kMaxInterfaces + 1];
  int n_itf = 0;
  double *h_pinned = nullptr;          // small pinned readback buffer
  int64_t *h_pinned_i = nullptr;
};

Runtime &rt();
Block *get_block(int id);
int ensure_init();

// > 0: the device allocation that many calls from now fails as if hipMalloc
// had (biocfd_inject_alloc_failure)
extern int g_alloc_fail_in;

template <typename T>
int dev_alloc(T **p, int64_t n) {
  if (*p) { (void)hipFree(*p); *p = nullptr; }
  if (n <= 0) n = 1;
  if (g_alloc_fail_in > 0 && --g_alloc_fail_in == 0) {
    set_error("hipMalloc(%lld bytes): injected allocation failure", (long long)(sizeof(T) * n));
    return -1;
  }
  hipError_t e = hipMalloc((void **)p, sizeof(T) * (size_t)n);
  if (e != hipSuccess) {
    set_error("hipMalloc(%lld bytes): %s", (long long)(sizeof(T) * n), hipGetErrorString(e));
    *p = nullptr;
    return -1;
  }
  return 0;
}

template <typename T>
int grow(T **p, int64_t *cap, int64_t need) {
  if (need <= *cap && *p) return 0;
  int64_t c = need < 64 ? 64 : need + need / 4;
  if (dev_alloc(p, c)) return -1;
  *cap = c;
  return 0;
}

inline unsigned div_up(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// b.f <-> b.ad.s for the seven fields while it lives: the tangent-linear
// passes run the forward routines, linear in the fields at frozen geometry,
// on the tangents
struct SwapFields {
  Block &b;
  explicit SwapFields(Block &blk) : b(blk) { flip(); }
  ~SwapFields() { flip(); }
  void flip() {
    for (int q = 0; q < 7; ++q) {
      double *t = b.f[q];
      b.f[q] = b.ad.s[q];
      b.ad.s[q] = t;
    }
  }
};

// internal entry points shared between translation units
int poisson_block_setup(Block &b);
int ib_list_capacity(Block &b, int64_t n_ib);
int build_cell_lists(Block &b, bool ib_lists);
int build_ts(Block &b);
int precision_mode();
int health_after_solve(int first, int last);
int stress_scratch_init(int device);
int mixed_sor(Block &b, int64_t pc_ita_max, double dt);
// tangent-linear (tlm) and adjoint (adj) passes over b.ad (geometry frozen)
int ad_ready(Block &b);
int ad_ghosts(Block &b);
int momentum_tlm(Block &b, double dt, double re);
int momentum_adj(Block &b, double dt, double re);
// rec: the record slot of the forward call whose values the geometric part
// reads (-1: geometry frozen)
int forcing_tlm(Block &b, int routine, const double kin[BK_N], const double kin_dot[BK_N], int rec);
int forcing_adj(Block &b, int routine, const double kin[BK_N], int rec);
int stress_tlm(Block &b, double mu_f, double rho_f, const double kin[BK_N], const double kin_dot[BK_N], int geom,
               double out[4]);
int stress_adj(Block &b, double mu_f, double rho_f, const double kin[BK_N], const double cbar[4], int geom);
int ad_geometry(Block &b);
void ad_release_records(Block &b);
int compute_div(Block &b);
int redblack_sor(Block &b, int64_t pc_ita_max, double dt, double epsi);
int correct(Block &b, double dt);
int velocity_bc(Block &b, double dt, double uc, double inlet);
int norm_copy(Block &b, double dt);
int projection(Block &b, int64_t n_iter, double dt);
int stress_sums(Block &b, double mu_f, double rho_f, const double kin[BK_N], double out[4]);
void ad_release(Block &b);
void sor_graphs_release(Block &b);
void batch_graphs_release();
int sor_graphs_on();
int mixed_sor_set(const std::vector<Block *> &blocks, int64_t pc_ita_max, double dt);
bool mixed_batchable();

}  // namespace bcfd
