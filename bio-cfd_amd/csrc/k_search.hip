// Immersed-boundary geometry, tagging and list compaction on gfx950
// (reference src/Search.F90: computeSurfaceNorm :210-258, tagging_th_core
// :260-386, selectiveRetagging_th :520-651, cellCount_solid :653-741,
// computeNormDistance :743-861, fine_block_cell :863-914, findTScells
// :434-518, cellCount_solid_coarse :1463-1532).
//
// Index lists come out of ordered (stable) scans over the interior in the
// reference's serial i-fastest order, so they are bit-identical to the CPU
// build instead of following atomic-capture order.  Nearest-element searches
// keep the reference's first-minimum tie-break (strict <, lowest index wins):
// the set-up tagging of the search box scans every element in order (LDS
// tiles of centroids); the per-step searches near the body (retagging,
// normal distances) go through a uniform grid of the centroids, one
// wavefront per cell (bin_search.hpp), with the same result.
#include "bin_search.hpp"
#include "biocfd_rt.hpp"

namespace bcfd {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kTile = kScanThreads * kScanItems;

// ---------------- ordered compaction over the block interior ----------------
struct ClsTag {  // 0 fluid, 1 solid, 2 IB
  const uint8_t *cell;
  __device__ int operator()(const Dims &d, int i, int j, int k) const {
    const uint8_t v = cell[ix3(d, i, j, k)];
    return v <= 2 ? (int)v : -1;
  }
};
struct ClsRedBlack {  // among fluid cells: 0 red ((i+j+k) odd), 1 black
  const uint8_t *cell;
  __device__ int operator()(const Dims &d, int i, int j, int k) const {
    if (cell[ix3(d, i, j, k)] != 0) return -1;
    return ((i + j + k) % 2 == 1) ? 0 : 1;
  }
};
struct ClsFluid {
  const uint8_t *cell;
  __device__ int operator()(const Dims &d, int i, int j, int k) const {
    return cell[ix3(d, i, j, k)] == 0 ? 0 : -1;
  }
};
struct ClsMask {
  const uint8_t *mask;
  __device__ int operator()(const Dims &d, int i, int j, int k) const {
    return mask[ix3(d, i, j, k)] ? 0 : -1;
  }
};

struct OutIJK {  // int32 (i list, j list, k list) per class, capacity cap[c]
  int *p[3];
  int64_t cap[3];
  int *ib_map;   // optional: map cell -> ordinal of class 2
  int map_class;
  __device__ void operator()(const Dims &d, int cls, int64_t r, int i, int j, int k) const {
    int *o = p[cls];
    if (!o) return;
    o[r] = i;
    o[r + cap[cls]] = j;
    o[r + 2 * cap[cls]] = k;
    if (ib_map && cls == map_class) ib_map[ix3(d, i, j, k)] = (int)r;
  }
};
struct OutLin {
  int *p;
  __device__ void operator()(const Dims &d, int cls, int64_t r, int i, int j, int k) const {
    p[r] = (i - 2) + d.nx * ((j - 2) + d.ny * (k - 2));
  }
};

__device__ inline void lin_to_ijk(const Dims &d, int64_t q, int &i, int &j, int &k) {
  i = 2 + (int)(q % d.nx);
  const int64_t r = q / d.nx;
  j = 2 + (int)(r % d.ny);
  k = 2 + (int)(r / d.ny);
}

// exclusive scan of one int per thread across the workgroup (kScanThreads)
__device__ inline int block_exscan(int v, int *s, int *total) {
  const int t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  for (int o = 1; o < kScanThreads; o <<= 1) {
    int a = (t >= o) ? s[t - o] : 0;
    __syncthreads();
    s[t] += a;
    __syncthreads();
  }
  const int incl = s[t];
  *total = s[kScanThreads - 1];
  __syncthreads();
  return incl - v;
}

template <class C>
__global__ void __launch_bounds__(kScanThreads) k_tile_count(Dims d, int64_t n_int, C cls, int ncls,
                                                             int64_t *tile_counts, int64_t ntiles) {
  __shared__ int s[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kScanItems;
  int cnt[3] = {0, 0, 0};
  for (int q = 0; q < kScanItems; ++q) {
    const int64_t lin = base + q;
    if (lin >= n_int) break;
    int i, j, k;
    lin_to_ijk(d, lin, i, j, k);
    const int c = cls(d, i, j, k);
    if (c >= 0) cnt[c]++;
  }
  for (int c = 0; c < ncls; ++c) {
    int tot;
    block_exscan(cnt[c], s, &tot);
    if (threadIdx.x == 0) tile_counts[(int64_t)c * ntiles + blockIdx.x] = tot;
  }
}

// exclusive scan of the tile counts of every class; totals[c] = class size
__global__ void __launch_bounds__(1024) k_scan_tiles(int64_t *tile_counts, int64_t ntiles, int ncls,
                                                     int64_t *totals) {
  __shared__ int64_t s[1024];
  const int t = threadIdx.x;
  const int64_t per = (ntiles + 1023) / 1024;
  for (int c = 0; c < ncls; ++c) {
    int64_t *tc = tile_counts + (int64_t)c * ntiles;
    int64_t sum = 0;
    for (int64_t q = t * per; q < (t + 1) * per && q < ntiles; ++q) sum += tc[q];
    s[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      int64_t a = (t >= o) ? s[t - o] : 0;
      __syncthreads();
      s[t] += a;
      __syncthreads();
    }
    int64_t run = s[t] - sum;
    for (int64_t q = t * per; q < (t + 1) * per && q < ntiles; ++q) {
      const int64_t v = tc[q];
      tc[q] = run;
      run += v;
    }
    if (t == 1023) totals[c] = s[1023];
    __syncthreads();
  }
}

template <class C, class O>
__global__ void __launch_bounds__(kScanThreads) k_tile_scatter(Dims d, int64_t n_int, C cls, int ncls,
                                                               const int64_t *tile_offsets, int64_t ntiles,
                                                               O out) {
  __shared__ int s[kScanThreads];
  const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kScanItems;
  int cnt[3] = {0, 0, 0};
  int cl[kScanItems];
  for (int q = 0; q < kScanItems; ++q) {
    const int64_t lin = base + q;
    cl[q] = -1;
    if (lin >= n_int) continue;
    int i, j, k;
    lin_to_ijk(d, lin, i, j, k);
    cl[q] = cls(d, i, j, k);
    if (cl[q] >= 0) cnt[cl[q]]++;
  }
  int64_t off[3];
  for (int c = 0; c < ncls; ++c) {
    int tot;
    const int ex = block_exscan(cnt[c], s, &tot);
    off[c] = tile_offsets[(int64_t)c * ntiles + blockIdx.x] + ex;
  }
  for (int q = 0; q < kScanItems; ++q) {
    if (cl[q] < 0) continue;
    int i, j, k;
    lin_to_ijk(d, base + q, i, j, k);
    out(d, cl[q], off[cl[q]]++, i, j, k);
  }
}

// two passes: count + scan (totals to host), then scatter with `make_out`
template <class C, class MakeOut>
static int ordered_compaction(Block &b, C cls, int ncls, int64_t totals_out[3], MakeOut make_out) {
  hipStream_t s = rt().stream;
  const int64_t n_int = (int64_t)b.d.nx * b.d.ny * b.d.nz;
  const int64_t ntiles = (n_int + kTile - 1) / kTile;
  if (grow(&b.scan_tmp, &b.scan_cap, ncls * ntiles + 8)) return -1;
  int64_t *totals = b.scan_tmp + ncls * ntiles;
  hipLaunchKernelGGL((k_tile_count<C>), dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, b.d, n_int, cls,
                     ncls, b.scan_tmp, ntiles);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, b.scan_tmp, ntiles, ncls, totals);
  BCFD_CHECK_LAUNCH();
  int64_t *h = rt().h_pinned_i;
  BCFD_HIP(hipMemcpyAsync(h, totals, sizeof(int64_t) * ncls, hipMemcpyDeviceToHost, s));
  BCFD_HIP(hipStreamSynchronize(s));
  for (int c = 0; c < ncls; ++c) totals_out[c] = h[c];
  auto out = make_out(totals_out);
  if (out.ok < 0) return -1;
  if (out.ok > 0) return 0;  // counts only
  hipLaunchKernelGGL((k_tile_scatter<C, decltype(out.o)>), dim3((unsigned)ntiles), dim3(kScanThreads), 0, s,
                     b.d, n_int, cls, ncls, b.scan_tmp, ntiles, out.o);
  BCFD_CHECK_LAUNCH();
  return 0;
}

template <class O>
struct OutSel {
  int ok;
  O o;
};

__global__ void k_unmap_ib(Dims d, const int *ijk, int64_t cap, int64_t n, int *ib_map) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) ib_map[ix3(d, ijk[q], ijk[q + cap], ijk[q + 2 * cap])] = -1;
}

int ib_list_capacity(Block &b, int64_t n_ib) {
  IbData &ib = b.ib;
  if (n_ib <= ib.cap_ib && ib.ib_ijk) return 0;
  const int64_t cap = n_ib + n_ib / 4 + 64;
  if (dev_alloc(&ib.ib_ijk, 3 * cap)) return -1;
  for (int q = 0; q < BN_NNEL; ++q)
    if (dev_alloc(&ib.nel[q], cap) || dev_alloc(&ib.normdis[q], cap)) return -1;
  ib.cap_ib = cap;
  return 0;
}

// cellCount_solid: device lists of solid and IB cells, counts of all classes,
// red/black counts; the fluid/red/black lists are materialised on demand
int build_cell_lists(Block &b, bool ib_lists) {
  hipStream_t s = rt().stream;
  IbData &ib = b.ib;
  // forget the previous IB ordinals
  if (ib_lists && ib.ib_ijk && ib.n_mapped > 0)
    hipLaunchKernelGGL(k_unmap_ib, dim3(div_up(ib.n_mapped, 256)), dim3(256), 0, s, b.d, ib.ib_ijk,
                       ib.cap_ib, ib.n_mapped, ib.ib_map);
  if (ib_lists) ib.n_mapped = 0;
  int64_t tot[3] = {0, 0, 0};
  int rc = ordered_compaction(b, ClsTag{b.tag[BT_CELL]}, 3, tot, [&](int64_t *t) {
    OutSel<OutIJK> r{0, {}};
    r.o.p[0] = nullptr;
    r.o.cap[0] = 0;
    r.o.p[1] = r.o.p[2] = nullptr;
    r.o.cap[1] = r.o.cap[2] = 0;
    r.o.ib_map = nullptr;
    r.o.map_class = 2;
    if (ib_lists) {
      if (grow(&ib.solid_ijk, &ib.cap_solid, 3 * (t[1] + 1))) { r.ok = -1; return r; }
      if (ib_list_capacity(b, t[2] + 1)) { r.ok = -1; return r; }
      r.o.p[1] = ib.solid_ijk;
      r.o.cap[1] = ib.cap_solid / 3;
      r.o.p[2] = ib.ib_ijk;
      r.o.cap[2] = ib.cap_ib;
      r.o.ib_map = ib.ib_map;
    }
    return r;
  });
  if (rc) return -1;
  b.counts[BC_FLUID] = tot[0];
  b.counts[BC_SOLID] = tot[1];
  b.counts[BC_IB] = tot[2];
  if (ib_lists) ib.n_mapped = tot[2];
  // red/black counts only (no scatter needed): count pass + scan
  int64_t rb[3] = {0, 0, 0};
  rc = ordered_compaction(b, ClsRedBlack{b.tag[BT_CELL]}, 2, rb, [&](int64_t *) {
    OutSel<OutIJK> r{1, {}};
    for (int c = 0; c < 3; ++c) { r.o.p[c] = nullptr; r.o.cap[c] = 0; }
    r.o.ib_map = nullptr;
    r.o.map_class = -1;
    return r;
  });
  if (rc) return -1;
  b.counts[BC_RED] = rb[0];
  b.counts[BC_BLACK] = rb[1];
  return 0;
}

int compact_mask_lin(Block &b, const uint8_t *mask, int **out, int64_t *cap, int64_t *count) {
  int64_t tot[3] = {0, 0, 0};
  int rc = ordered_compaction(b, ClsMask{mask}, 1, tot, [&](int64_t *t) {
    OutSel<OutLin> r{0, {nullptr}};
    if (grow(out, cap, t[0] + 1)) { r.ok = -1; return r; }
    r.o.p = *out;
    return r;
  });
  *count = tot[0];
  return rc;
}

// ---------------- geometry ----------------
__global__ void k_surface_norm(int64_t ne, const int *e1, const int *e2, const int *e3, const double *x,
                               const double *y, const double *z, double *xc, double *yc, double *zc,
                               double *ca, double *cb, double *cg, double *len, double inor) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= ne) return;
  const int a = e1[n] - 1, b = e2[n] - 1, c = e3[n] - 1;
  const double p1x = x[a], p1y = y[a], p1z = z[a];
  const double p2x = x[b], p2y = y[b], p2z = z[b];
  const double p3x = x[c], p3y = y[c], p3z = z[c];
  xc[n] = (p2x + p1x + p3x) / 3.0;
  yc[n] = (p2y + p1y + p3y) / 3.0;
  zc[n] = (p2z + p1z + p3z) / 3.0;
  const double nxv = (p2y - p1y) * (p3z - p1z) - (p3y - p1y) * (p2z - p1z);
  const double nyv = (p2z - p1z) * (p3x - p1x) - (p3z - p1z) * (p2x - p1x);
  const double nzv = (p2x - p1x) * (p3y - p1y) - (p3x - p1x) * (p2y - p1y);
  const double l = sqrt(nxv * nxv + nyv * nyv + nzv * nzv);
  len[n] = l;
  ca[n] = inor * nxv / l;
  cb[n] = inor * nyv / l;
  cg[n] = inor * nzv / l;
}

constexpr int kElemTile = 256;

struct Geo {
  Dims d;
  const double *x1, *y1, *z1, *xp, *yp, *zp;
  const double *xc, *yc, *zc, *ca, *cb, *cg;
  int64_t ne;
};

// nearest-centroid searches for the cell centre and the cell's low corner;
// all threads of the workgroup must call it (LDS tiles of centroids)
__device__ void nearest_pair(const Geo &g, bool active, double xpi, double ypj, double zpk, double x1i,
                             double y1j, double z1k, int &ncen, int &npnt, double *sx, double *sy, double *sz) {
  double md = 1e14, md1 = 1e14;
  int best = 0, best1 = 0;
  for (int64_t base = 0; base < g.ne; base += kElemTile) {
    const int nt = (int)((g.ne - base) < kElemTile ? (g.ne - base) : kElemTile);
    __syncthreads();
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
      sx[t] = g.xc[base + t];
      sy[t] = g.yc[base + t];
      sz[t] = g.zc[base + t];
    }
    __syncthreads();
    if (active) {
      for (int t = 0; t < nt; ++t) {
        const double ax = xpi - sx[t], ay = ypj - sy[t], az = zpk - sz[t];
        const double dc = ax * ax + ay * ay + az * az;
        const double bx = x1i - sx[t], by = y1j - sy[t], bz = z1k - sz[t];
        const double dp = bx * bx + by * by + bz * bz;
        if (dc < md) { md = dc; best = (int)(base + t) + 1; }
        if (dp < md1) { md1 = dp; best1 = (int)(base + t) + 1; }
      }
    }
  }
  ncen = best;
  npnt = best1;
}

// tag one cell from its nearest elements (Search.F90:305-338)
__device__ void tag_from(const Geo &g, uint8_t *cell, uint8_t *node_tag, int i, int j, int k, int ncen, int npnt) {
  const int e = ncen - 1;
  if ((g.x1[i - 1] <= g.xc[e] && g.x1[i] >= g.xc[e]) && (g.y1[j - 1] <= g.yc[e] && g.y1[j] >= g.yc[e]) &&
      (g.z1[k - 1] <= g.zc[e] && g.z1[k] >= g.zc[e]))
    cell[ix3(g.d, i, j, k)] = 2;
  const int f = npnt - 1;
  const double n2dotn = (g.x1[i - 1] - g.xc[f]) * g.ca[f] + (g.y1[j - 1] - g.yc[f]) * g.cb[f] +
                        (g.z1[k - 1] - g.zc[f]) * g.cg[f];
  node_tag[ixt(g.d, i, j, k)] = (n2dotn >= -1e-16) ? 0 : 1;
}

__global__ void __launch_bounds__(256) k_tag_box(Geo g, int i0, int ni, int j0, int nj, int k0, int nk,
                                                 uint8_t *cell, uint8_t *node_tag) {
  __shared__ double sx[kElemTile], sy[kElemTile], sz[kElemTile];
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = q < (int64_t)ni * nj * nk;
  int i = i0, j = j0, k = k0;
  if (active) {
    i = i0 + (int)(q % ni);
    j = j0 + (int)((q / ni) % nj);
    k = k0 + (int)(q / ((int64_t)ni * nj));
  }
  int nc, np;
  nearest_pair(g, active, g.xp[i - 1], g.yp[j - 1], g.zp[k - 1], g.x1[i - 1], g.y1[j - 1], g.z1[k - 1], nc, np,
               sx, sy, sz);
  if (active) tag_from(g, cell, node_tag, i, j, k, nc, np);
}

__device__ inline int corner_sum(const Dims &d, const uint8_t *t, int i, int j, int k) {
  return t[ixt(d, i, j, k)] + t[ixt(d, i + 1, j, k)] + t[ixt(d, i, j + 1, k)] + t[ixt(d, i + 1, j + 1, k)] +
         t[ixt(d, i, j, k + 1)] + t[ixt(d, i + 1, j, k + 1)] + t[ixt(d, i, j + 1, k + 1)] +
         t[ixt(d, i + 1, j + 1, k + 1)];
}

__global__ void k_solid_box(Dims d, int i0, int ni, int j0, int nj, int k0, int nk, uint8_t *cell,
                            const uint8_t *node_tag) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)ni * nj * nk) return;
  const int i = i0 + (int)(q % ni), j = j0 + (int)((q / ni) % nj), k = k0 + (int)(q / ((int64_t)ni * nj));
  const int64_t c = ix3(d, i, j, k);
  if (cell[c] != 2 && corner_sum(d, node_tag, i, j, k) == 8) cell[c] = 1;
}

// selectiveRetagging_th: previous IB cells -> 0, re-tag their 3x3x3
// neighbourhoods, re-classify (Search.F90:528-622)
__global__ void k_reset_ib(Dims d, const int *ijk, int64_t cap, int64_t n, uint8_t *cell) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) cell[ix3(d, ijk[q], ijk[q + cap], ijk[q + 2 * cap])] = 0;
}

// ---- uniform grid of the centroids ----
// bounding box of the centroids -> bin grid; bin counters to zero
__global__ void __launch_bounds__(1024) k_bin_setup(const double *xc, const double *yc, const double *zc,
                                                    int64_t ne, BinGrid *bg, int *fill) {
  __shared__ double s_mn[3][16], s_mx[3][16];
  double mn[3] = {1e308, 1e308, 1e308}, mx[3] = {-1e308, -1e308, -1e308};
  for (int64_t e = threadIdx.x; e < ne; e += blockDim.x) {
    const double v[3] = {xc[e], yc[e], zc[e]};
    for (int a = 0; a < 3; ++a) {
      mn[a] = v[a] < mn[a] ? v[a] : mn[a];
      mx[a] = v[a] > mx[a] ? v[a] : mx[a];
    }
  }
  for (int a = 0; a < 3; ++a)
    for (int off = 32; off >= 1; off >>= 1) {
      const double o1 = __shfl_xor(mn[a], off), o2 = __shfl_xor(mx[a], off);
      mn[a] = o1 < mn[a] ? o1 : mn[a];
      mx[a] = o2 > mx[a] ? o2 : mx[a];
    }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; ++a) { s_mn[a][w] = mn[a]; s_mx[a][w] = mx[a]; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q)
      for (int a = 0; a < 3; ++a) {
        mn[a] = s_mn[a][q] < mn[a] ? s_mn[a][q] : mn[a];
        mx[a] = s_mx[a][q] > mx[a] ? s_mx[a][q] : mx[a];
      }
    *bg = bin_grid_of(mn, mx);
  }
  for (int q = threadIdx.x; q < kMaxBins; q += blockDim.x) fill[q] = 0;
}

__global__ void k_bin_count(const double *xc, const double *yc, const double *zc, int64_t ne, const BinGrid *bg,
                            int *fill) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < ne) atomicAdd(&fill[bin_of(*bg, xc[e], yc[e], zc[e])], 1);
}

// exclusive scan of the bin counts into start[0..nb]; the counters become
// the bins' fill cursors
__global__ void __launch_bounds__(1024) k_bin_scan(const BinGrid *bg, int *fill, int *start) {
  __shared__ int s_sum[1024];
  const int nb = bg->nx * bg->ny * bg->nz;
  const int per = (nb + (int)blockDim.x - 1) / (int)blockDim.x;
  const int b0 = threadIdx.x * per, b1 = b0 + per < nb ? b0 + per : nb;
  int sum = 0;
  for (int b = b0; b < b1; ++b) sum += fill[b];
  s_sum[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    const int v = threadIdx.x >= (unsigned)off ? s_sum[threadIdx.x - off] : 0;
    __syncthreads();
    s_sum[threadIdx.x] += v;
    __syncthreads();
  }
  int run = s_sum[threadIdx.x] - sum;  // exclusive prefix of this thread's chunk
  for (int b = b0; b < b1; ++b) {
    const int c = fill[b];
    start[b] = run;
    fill[b] = run;
    run += c;
  }
  if (threadIdx.x == blockDim.x - 1) start[nb] = s_sum[threadIdx.x];
}

__global__ void k_bin_fill(const double *xc, const double *yc, const double *zc, int64_t ne, const BinGrid *bg,
                           int *fill, int *belem) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < ne) belem[atomicAdd(&fill[bin_of(*bg, xc[e], yc[e], zc[e])], 1)] = (int)e;
}

struct BinView {
  const BinGrid *bg;
  const int *start, *elem;
};

// lexicographic (distance, element) minimum over the wavefront
__device__ inline void wave_lexmin(double &d, int &e) {
  for (int off = 32; off >= 1; off >>= 1) {
    const double od = __shfl_xor(d, off);
    const int oe = __shfl_xor(e, off);
    if (lex_less(od, oe, d, e)) {
      d = od;
      e = oe;
    }
  }
}

// nearest elements (1-based, 0 for none) of T points of one cell, by the
// calling wavefront; every lane gets the results
template <int T>
__device__ void wave_nearest(const Geo &g, const BinView &bv, const double *px, const double *py, const double *pz,
                             int *out) {
  const BinGrid bg = *bv.bg;
  const NearestQuery<T> Q = make_query<T>(bg, px, py, pz);
  const int lane = threadIdx.x & 63;
  double bd[T];
  int be[T];
  for (int s = 0; s < T; ++s) { bd[s] = kNoDist; be[s] = kNoElem; }
  for (int r = 0;; ++r) {
    shell_scan<T>(bg, bv.start, bv.elem, g.xc, g.yc, g.zc, Q, r, lane, kSearchLanes, bd, be);
    for (int s = 0; s < T; ++s) wave_lexmin(bd[s], be[s]);
    if (shell_done<T>(bg, Q, r, bd)) break;
  }
  for (int s = 0; s < T; ++s) out[s] = nearest_result(bd[s], be[s]);
}

// selectiveRetagging_th's re-tag of the 3x3x3 neighbourhoods (Search.F90:
// 540-560 with tagging_th_core :305-338): one wavefront per neighbour entry.
// Neighbourhoods of adjacent IB cells overlap; the tags of a cell depend on
// its position and the surface only, so the reference's repeated visits
// write the same values.  Here the first wavefront to claim a cell's bit
// tags it and the others return; k_reclass clears the bits again.
__global__ void __launch_bounds__(256) k_retag(Geo g, BinView bv, const int *ijk, int64_t cap, int64_t n,
                                               uint8_t *cell, uint8_t *node_tag, unsigned *bits) {
  const int64_t q = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (q >= n * 27) return;  // the whole wavefront
  const int64_t nn = q / 27;
  const int r = (int)(q % 27);
  const int i = ijk[nn] + (r % 3) - 1, j = ijk[nn + cap] + ((r / 3) % 3) - 1, k = ijk[nn + 2 * cap] + (r / 9) - 1;
  const int64_t c = ix3(g.d, i, j, k);
  unsigned taken = 0;
  if ((threadIdx.x & 63) == 0) {
    const unsigned bit = 1u << (unsigned)(c & 31);
    taken = atomicOr(&bits[c >> 5], bit) & bit;
  }
  if (__shfl(taken, 0, 64)) return;  // the whole wavefront: another one tags this cell
  const double px[2] = {g.xp[i - 1], g.x1[i - 1]}, py[2] = {g.yp[j - 1], g.y1[j - 1]},
               pz[2] = {g.zp[k - 1], g.z1[k - 1]};
  int best[2];
  wave_nearest<2>(g, bv, px, py, pz, best);
  if ((threadIdx.x & 63) == 0) tag_from(g, cell, node_tag, i, j, k, best[0], best[1]);
}

__global__ void k_reclass(Dims d, const int *ijk, int64_t cap, int64_t n, uint8_t *cell, const uint8_t *node_tag,
                          unsigned *bits) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * 27) return;
  const int64_t nn = q / 27;
  const int r = (int)(q % 27);
  const int i = ijk[nn] + (r % 3) - 1, j = ijk[nn + cap] + ((r / 3) % 3) - 1, k = ijk[nn + 2 * cap] + (r / 9) - 1;
  const int64_t c = ix3(d, i, j, k);
  if (cell[c] != 2) cell[c] = (corner_sum(d, node_tag, i, j, k) == 8) ? 1 : 0;
  atomicAnd(&bits[c >> 5], ~(1u << (unsigned)(c & 31)));
}

// computeNormDistance: seven nearest-element searches per IB cell and the
// signed normal distances (Search.F90:762-858)
__global__ void __launch_bounds__(256) k_norm_distance(Geo g, BinView bv, const int *ijk, int64_t cap, int64_t n,
                                                       int *nel0, int *nel1, int *nel2, int *nel3, int *nel4,
                                                       int *nel5, int *nel6, double *nd0, double *nd1,
                                                       double *nd2, double *nd3, double *nd4, double *nd5,
                                                       double *nd6) {
  const int64_t q = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (q >= n) return;  // the whole wavefront
  const int i = ijk[q], j = ijk[q + cap], k = ijk[q + 2 * cap];
  const double n1x = g.xp[i - 1], n2x = g.x1[i - 1], n3x = g.x1[i];
  const double n1y = g.yp[j - 1], n2y = g.y1[j - 1], n3y = g.y1[j];
  const double n1z = g.zp[k - 1], n2z = g.z1[k - 1], n3z = g.z1[k];
  // the p point and the six face points (u: x faces, v: y faces, w: z faces)
  const double px[7] = {n1x, n2x, n3x, n1x, n1x, n1x, n1x};
  const double py[7] = {n1y, n1y, n1y, n2y, n3y, n1y, n1y};
  const double pz[7] = {n1z, n1z, n1z, n1z, n1z, n2z, n3z};
  int mi[7];
  wave_nearest<7>(g, bv, px, py, pz, mi);
  const int lane = threadIdx.x & 63;
  if (lane >= 7) return;
  int *nel[7] = {nel0, nel1, nel2, nel3, nel4, nel5, nel6};
  double *nd[7] = {nd0, nd1, nd2, nd3, nd4, nd5, nd6};
  const int s = lane, e = mi[s] - 1;
  nel[s][q] = mi[s];
  nd[s][q] = e >= 0 ? (px[s] - g.xc[e]) * g.ca[e] + (py[s] - g.yc[e]) * g.cb[e] + (pz[s] - g.zc[e]) * g.cg[e] : 0.0;
}

// findTScells (Search.F90:434-518)
__global__ void k_clear_cell2(Dims d, uint8_t *cell2) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (i > d.nx + 1 || j > d.ny + 1) return;
  cell2[ixt(d, i, j, k)] = 0;
}

__global__ void __launch_bounds__(1024) k_ts_build(Dims d, const int *ijk, int64_t cap, int64_t n_ib,
                                                   const int *nelp, const int *surf, uint8_t *cell2, int *ts_ijk,
                                                   int64_t cap_ts, int *index_ts, int64_t *count) {
  __shared__ int s[1024];
  __shared__ int64_t s_base;
  if (threadIdx.x == 0) s_base = 0;
  __syncthreads();
  for (int64_t base = 0; base < n_ib; base += 1024) {
    const int64_t q = base + threadIdx.x;
    int f = 0;
    if (q < n_ib) {
      const int sid = surf[nelp[q] - 1];
      f = (sid == 51 || sid == 52) ? 1 : 0;
    }
    s[threadIdx.x] = f;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      int a = (threadIdx.x >= (unsigned)o) ? s[threadIdx.x - o] : 0;
      __syncthreads();
      s[threadIdx.x] += a;
      __syncthreads();
    }
    if (f) {
      const int64_t r = s_base + s[threadIdx.x] - 1;
      const int i = ijk[q], j = ijk[q + cap], k = ijk[q + 2 * cap];
      cell2[ixt(d, i, j, k)] = 2;
      if (ts_ijk) {
        ts_ijk[r] = i; ts_ijk[r + cap_ts] = j; ts_ijk[r + 2 * cap_ts] = k;
        index_ts[r] = (int)q + 1;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_base += s[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = s_base;
}

// fine_block_cell (Search.F90:863-914)
__global__ void k_fine_block_cell(Dims d, const double *xp, const double *yp, const double *zp, double xs,
                                  double xe, double ys, double ye, double zs, double ze, double mx, double my,
                                  double mz, uint8_t *cell, uint8_t *cell_n, uint8_t *cell_pr) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 2 + blockIdx.z;
  if (i > d.nx + 1 || j > d.ny + 1) return;
  const double x = xp[i - 1], y = yp[j - 1], z = zp[k - 1];
  if (x >= xs + mx && x <= xe - mx && z >= zs + mz && z <= ze - mz && y >= ys + my && y <= ye - my) {
    cell_n[ix3(d, i, j, k)] = 1;
    cell[ix3(d, i, j, k)] = 1;
  }
  if (x >= xs && x <= xe && z >= zs && z <= ze && y >= ys && y <= ye) cell_pr[ixt(d, i, j, k)] = 1;
}

static Geo geo_of(Block &b) {
  Geo g;
  g.d = b.d;
  g.x1 = b.g[BG_X1]; g.y1 = b.g[BG_Y1]; g.z1 = b.g[BG_Z1];
  g.xp = b.g[BG_XP]; g.yp = b.g[BG_YP]; g.zp = b.g[BG_ZP];
  g.xc = b.ib.elem[BE_XCENT]; g.yc = b.ib.elem[BE_YCENT]; g.zc = b.ib.elem[BE_ZCENT];
  g.ca = b.ib.elem[BE_COSA]; g.cb = b.ib.elem[BE_COSB]; g.cg = b.ib.elem[BE_COSG];
  g.ne = b.ib.n_elems;
  return g;
}

// the bin grid of the current centroids (after k_surface_norm)
static int build_bins(Block &b) {
  IbData &ib = b.ib;
  hipStream_t s = rt().stream;
  if (!ib.bin_grid && dev_alloc((char **)&ib.bin_grid, (int64_t)sizeof(BinGrid))) return -1;
  if (!ib.bin_start && dev_alloc(&ib.bin_start, kMaxBins + 1)) return -1;
  if (!ib.bin_fill && dev_alloc(&ib.bin_fill, kMaxBins)) return -1;
  if (!ib.bin_elem && dev_alloc(&ib.bin_elem, ib.n_elems)) return -1;
  const double *xc = ib.elem[BE_XCENT], *yc = ib.elem[BE_YCENT], *zc = ib.elem[BE_ZCENT];
  BinGrid *bg = (BinGrid *)ib.bin_grid;
  const dim3 ge(div_up(ib.n_elems, 256));
  hipLaunchKernelGGL(k_bin_setup, dim3(1), dim3(1024), 0, s, xc, yc, zc, ib.n_elems, bg, ib.bin_fill);
  hipLaunchKernelGGL(k_bin_count, ge, dim3(256), 0, s, xc, yc, zc, ib.n_elems, bg, ib.bin_fill);
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, bg, ib.bin_fill, ib.bin_start);
  hipLaunchKernelGGL(k_bin_fill, ge, dim3(256), 0, s, xc, yc, zc, ib.n_elems, bg, ib.bin_fill, ib.bin_elem);
  BCFD_CHECK_LAUNCH();
  ib.bins_ready = true;
  return 0;
}

static BinView bins_of(const Block &b) {
  return BinView{(const BinGrid *)b.ib.bin_grid, b.ib.bin_start, b.ib.bin_elem};
}

static Block *ib_block(int id) {
  Block *b = get_block(id);
  if (!b) return nullptr;
  if (!b->ib.present) { set_error("block %d has no immersed body", id); return nullptr; }
  for (int q : {BG_X1, BG_Y1, BG_Z1, BG_XP, BG_YP, BG_ZP})
    if (!b->g[q]) { set_error("block %d grids not uploaded", id); return nullptr; }
  return b;
}

int build_ts(Block &b) {
  hipStream_t s = rt().stream;
  IbData &ib = b.ib;
  dim3 blk(64, 4, 1);
  hipLaunchKernelGGL(k_clear_cell2, dim3(div_up(b.d.nx, 64), div_up(b.d.ny, 4), b.d.nz), blk, 0, s, b.d,
                     b.tag[BT_CELL2]);
  const int64_t n_ib = b.counts[BC_IB];
  const int64_t need = n_ib + 1;
  if (need > ib.cap_ts || !ib.ts_ijk) {
    const int64_t cap = need + need / 4 + 64;
    if (dev_alloc(&ib.ts_ijk, 3 * cap) || dev_alloc(&ib.index_ts, cap)) return -1;
    for (int q = 0; q < BGH_N; ++q)
      if (dev_alloc(&ib.ghost[q], cap)) return -1;
    ib.cap_ts = cap;
  }
  int64_t *cnt = b.scan_tmp;
  if (!cnt && grow(&b.scan_tmp, &b.scan_cap, 8)) return -1;
  cnt = b.scan_tmp;
  hipLaunchKernelGGL(k_ts_build, dim3(1), dim3(1024), 0, s, b.d, ib.ib_ijk, ib.cap_ib, n_ib, ib.nel[BN_P],
                     ib.surf, b.tag[BT_CELL2], ib.ts_ijk, ib.cap_ts, ib.index_ts, cnt);
  BCFD_CHECK_LAUNCH();
  BCFD_HIP(hipMemcpyAsync(rt().h_pinned_i, cnt, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  BCFD_HIP(hipStreamSynchronize(s));
  b.counts[BC_TS] = rt().h_pinned_i[0];
  for (int q = 0; q < BGH_N; ++q)
    BCFD_HIP(hipMemsetAsync(ib.ghost[q], 0, sizeof(double) * ib.cap_ts, s));
  return 0;
}

// ---- list downloads ----------------------------------------------------------
__global__ void k_widen_list(const int *src, int64_t cap_src, int64_t n, int64_t *dst) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  dst[q] = src[q];
  dst[q + n] = src[q + cap_src];
  dst[q + 2 * n] = src[q + 2 * cap_src];
}

static int download_ijk(const int *src, int64_t cap, int64_t n, int64_t *host) {
  if (n <= 0) return 0;
  int64_t *tmp = nullptr;
  if (dev_alloc(&tmp, 3 * n)) return -1;
  hipLaunchKernelGGL(k_widen_list, dim3(div_up(n, 256)), dim3(256), 0, rt().stream, src, cap, n, tmp);
  BCFD_CHECK_LAUNCH();
  BCFD_HIP(hipMemcpyAsync(host, tmp, sizeof(int64_t) * 3 * n, hipMemcpyDeviceToHost, rt().stream));
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  (void)hipFree(tmp);
  return 0;
}

template <class T, class U>
static int download_vec(const T *src, int64_t n, U *host) {
  if (n <= 0) return 0;
  std::vector<T> t((size_t)n);
  BCFD_HIP(hipMemcpyAsync(t.data(), src, sizeof(T) * n, hipMemcpyDeviceToHost, rt().stream));
  BCFD_HIP(hipStreamSynchronize(rt().stream));
  for (int64_t q = 0; q < n; ++q) host[q] = (U)t[q];
  return 0;
}

}  // namespace bcfd

using namespace bcfd;

extern "C" {

int biocfd_compute_surface_norm(int id, const double *x1, const double *y1, const double *z1, int64_t inor) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  IbData &ib = b->ib;
  if (!ib.present) { set_error("block %d has no immersed body", id); return -1; }
  hipStream_t s = rt().stream;
  const double *src[3] = {x1, y1, z1};
  for (int c = 0; c < 3; ++c)
    BCFD_HIP(hipMemcpyAsync(ib.node[c], src[c], sizeof(double) * ib.n_nodes, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_surface_norm, dim3(div_up(ib.n_elems, 256)), dim3(256), 0, s, ib.n_elems, ib.el[0],
                     ib.el[1], ib.el[2], ib.node[0], ib.node[1], ib.node[2], ib.elem[BE_XCENT],
                     ib.elem[BE_YCENT], ib.elem[BE_ZCENT], ib.elem[BE_COSA], ib.elem[BE_COSB],
                     ib.elem[BE_COSG], ib.elem[BE_ELEN], (double)inor);
  BCFD_CHECK_LAUNCH();
  if (build_bins(*b)) return -1;
  BCFD_HIP(hipStreamSynchronize(s));  // host node buffers may be reused by the caller
  return 0;
}

int biocfd_tagging(int id) {
  BCFD_API_LOCK();
  Block *b = ib_block(id);
  if (!b) return -1;
  hipStream_t s = rt().stream;
  BCFD_HIP(hipMemsetAsync(b->tag[BT_CELL], 0, b->n2, s));
  BCFD_HIP(hipMemsetAsync(b->tag[BT_CELL2], 0, b->n3, s));
  BCFD_HIP(hipMemsetAsync(b->tag[BT_NODEIDTAG], 0, b->n3, s));
  const int *sb = b->ib.search;
  const int ni = sb[1] - sb[0] + 1, nj = sb[3] - sb[2] + 1, nk = sb[5] - sb[4] + 1;
  const int64_t n = (int64_t)ni * nj * nk;
  Geo g = geo_of(*b);
  hipLaunchKernelGGL(k_tag_box, dim3(div_up(n, 256)), dim3(256), 0, s, g, sb[0], ni, sb[2], nj, sb[4], nk,
                     b->tag[BT_CELL], b->tag[BT_NODEIDTAG]);
  hipLaunchKernelGGL(k_solid_box, dim3(div_up(n, 256)), dim3(256), 0, s, b->d, sb[0], ni, sb[2], nj, sb[4], nk,
                     b->tag[BT_CELL], b->tag[BT_NODEIDTAG]);
  BCFD_CHECK_LAUNCH();
  // counts of the three classes (Search.F90:365-380)
  return build_cell_lists(*b, false);
}

int biocfd_selective_retagging(int id, int64_t blk_mv_tag) {
  BCFD_API_LOCK();
  Block *b = ib_block(id);
  if (!b) return -1;
  if (blk_mv_tag != 0) return 0;
  hipStream_t s = rt().stream;
  IbData &ib = b->ib;
  const int64_t n = ib.n_mapped;
  if (n > 0) {
    Geo g = geo_of(*b);
    hipLaunchKernelGGL(k_reset_ib, dim3(div_up(n, 256)), dim3(256), 0, s, b->d, ib.ib_ijk, ib.cap_ib, n,
                       b->tag[BT_CELL]);
    if (!ib.bins_ready) { set_error("selective_retagging: surface not computed"); return -1; }
    const int64_t words = div_up(b->n2, 32);
    if (ib.retag_words < words) {
      if (dev_alloc(&ib.retag_bits, words)) return -1;
      ib.retag_words = words;
      BCFD_HIP(hipMemsetAsync(ib.retag_bits, 0, sizeof(unsigned) * (size_t)words, s));
    }
    hipLaunchKernelGGL(k_retag, dim3(div_up(n * 27, 4)), dim3(256), 0, s, g, bins_of(*b), ib.ib_ijk, ib.cap_ib, n,
                       b->tag[BT_CELL], b->tag[BT_NODEIDTAG], ib.retag_bits);
    hipLaunchKernelGGL(k_reclass, dim3(div_up(n * 27, 256)), dim3(256), 0, s, b->d, ib.ib_ijk, ib.cap_ib, n,
                       b->tag[BT_CELL], b->tag[BT_NODEIDTAG], ib.retag_bits);
    BCFD_CHECK_LAUNCH();
  }
  // new counts (Search.F90:623-647); the lists are rebuilt by cell_count_solid
  int64_t tot[3] = {0, 0, 0};
  auto none = [&](int64_t *) {
    OutSel<OutIJK> r{1, {}};
    for (int c = 0; c < 3; ++c) { r.o.p[c] = nullptr; r.o.cap[c] = 0; }
    r.o.ib_map = nullptr;
    r.o.map_class = -1;
    return r;
  };
  if (ordered_compaction(*b, ClsTag{b->tag[BT_CELL]}, 3, tot, none)) return -1;
  b->counts[BC_FLUID] = tot[0];
  b->counts[BC_SOLID] = tot[1];
  b->counts[BC_IB] = tot[2];
  return 0;
}

int biocfd_cell_count_solid(int id) {
  BCFD_API_LOCK();
  Block *b = ib_block(id);
  if (!b) return -1;
  return build_cell_lists(*b, true);
}

int biocfd_cell_count_solid_coarse(int id) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  // only the fluid/red/black counts change (Search.F90:1468-1506)
  const int64_t solid = b->counts[BC_SOLID], ib = b->counts[BC_IB];
  if (build_cell_lists(*b, false)) return -1;
  b->counts[BC_SOLID] = solid;
  b->counts[BC_IB] = ib;
  return 0;
}

int biocfd_fine_block_cell(int coarse) {
  BCFD_API_LOCK();
  Block *a = get_block(coarse);
  if (!a) return -1;
  hipStream_t s = rt().stream;
  BCFD_HIP(hipMemsetAsync(a->tag[BT_CELL_N], 0, a->n2, s));
  BCFD_HIP(hipMemsetAsync(a->tag[BT_CELL], 0, a->n2, s));
  BCFD_HIP(hipMemsetAsync(a->tag[BT_CELL_PR], 0, a->n3, s));
  dim3 blk(64, 4, 1);
  dim3 grd(div_up(a->d.nx, 64), div_up(a->d.ny, 4), a->d.nz);
  for (int q = 1; q <= kMaxInterfaces; ++q) {
    Interface &it = rt().itf[q];
    if (!it.live || it.a != coarse) continue;
    Block *fb = get_block(it.b);
    if (!fb) return -1;
    const double ci = (double)fb->cintp;
    hipLaunchKernelGGL(k_fine_block_cell, grd, blk, 0, s, a->d, a->g[BG_XP], a->g[BG_YP], a->g[BG_ZP],
                       it.ext[0], it.ext[1], it.ext[2], it.ext[3], it.ext[4], it.ext[5], ci * a->dx, ci * a->dy,
                       ci * a->dz, a->tag[BT_CELL], a->tag[BT_CELL_N], a->tag[BT_CELL_PR]);
  }
  BCFD_CHECK_LAUNCH();
  return 0;
}

int biocfd_compute_norm_distance(int id) {
  BCFD_API_LOCK();
  Block *b = ib_block(id);
  if (!b) return -1;
  IbData &ib = b->ib;
  const int64_t n = b->counts[BC_IB];
  if (n == 0) return 0;
  if (!ib.ib_ijk) { set_error("compute_norm_distance: no IB list"); return -1; }
  if (!ib.bins_ready) { set_error("compute_norm_distance: surface not computed"); return -1; }
  Geo g = geo_of(*b);
  hipLaunchKernelGGL(k_norm_distance, dim3(div_up(n, 4)), dim3(256), 0, rt().stream, g, bins_of(*b), ib.ib_ijk,
                     ib.cap_ib, n, ib.nel[BN_P], ib.nel[BN_U1], ib.nel[BN_U2], ib.nel[BN_V1], ib.nel[BN_V2], ib.nel[BN_W1],
                     ib.nel[BN_W2], ib.normdis[BN_P], ib.normdis[BN_U1], ib.normdis[BN_U2], ib.normdis[BN_V1],
                     ib.normdis[BN_V2], ib.normdis[BN_W1], ib.normdis[BN_W2]);
  BCFD_CHECK_LAUNCH();
  return 0;
}

int biocfd_find_ts_cells(int id) {
  BCFD_API_LOCK();
  Block *b = ib_block(id);
  if (!b) return -1;
  if (b->counts[BC_IB] > 0 && !b->ib.ib_ijk) { set_error("find_ts_cells: no IB list"); return -1; }
  return build_ts(*b);
}

int biocfd_list_download(int id, int which, int64_t *host, int64_t count) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  IbData &ib = b->ib;
  const int cidx[BL_NLIST] = {BC_FLUID, BC_RED, BC_BLACK, BC_SOLID, BC_IB, BC_TS};
  if (which < 0 || which >= BL_NLIST) { set_error("bad list %d", which); return -1; }
  if (which == BL_RED || which == BL_BLACK || which == BL_FLUID) {
    int64_t tot[3] = {0, 0, 0};
    int *tmp = nullptr;
    int64_t cap = 0;
    int rc;
    if (which == BL_FLUID) {
      rc = ordered_compaction(*b, ClsFluid{b->tag[BT_CELL]}, 1, tot, [&](int64_t *t) {
        OutSel<OutIJK> r{0, {}};
        for (int c = 0; c < 3; ++c) { r.o.p[c] = nullptr; r.o.cap[c] = 0; }
        r.o.ib_map = nullptr;
        r.o.map_class = -1;
        if (grow(&tmp, &cap, 3 * (t[0] + 1))) { r.ok = -1; return r; }
        r.o.p[0] = tmp;
        r.o.cap[0] = t[0] + 1;
        return r;
      });
    } else {
      const int want = (which == BL_RED) ? 0 : 1;
      rc = ordered_compaction(*b, ClsRedBlack{b->tag[BT_CELL]}, 2, tot, [&](int64_t *t) {
        OutSel<OutIJK> r{0, {}};
        for (int c = 0; c < 3; ++c) { r.o.p[c] = nullptr; r.o.cap[c] = 0; }
        r.o.ib_map = nullptr;
        r.o.map_class = -1;
        if (grow(&tmp, &cap, 3 * (t[want] + 1))) { r.ok = -1; return r; }
        r.o.p[want] = tmp;
        r.o.cap[want] = t[want] + 1;
        return r;
      });
    }
    if (rc) return -1;
    const int64_t n = (which == BL_BLACK) ? tot[1] : tot[0];
    if (n != count) { set_error("list %d has %lld entries, caller expects %lld", which, (long long)n, (long long)count); (void)hipFree(tmp); return -1; }
    rc = download_ijk(tmp, n + 1, n, host);
    (void)hipFree(tmp);
    return rc;
  }
  const int64_t n = b->counts[cidx[which]];
  if (n != count) { set_error("list %d has %lld entries, caller expects %lld", which, (long long)n, (long long)count); return -1; }
  if (which == BL_SOLID) return download_ijk(ib.solid_ijk, ib.cap_solid / 3, n, host);
  if (which == BL_INTERCEPTED) return download_ijk(ib.ib_ijk, ib.cap_ib, n, host);
  return download_ijk(ib.ts_ijk, ib.cap_ts, n, host);
}

int biocfd_nel_download(int id, int which, int64_t *host, int64_t count) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (which < 0 || which >= BN_NNEL || count != b->counts[BC_IB]) { set_error("nel_download: bad request"); return -1; }
  return download_vec(b->ib.nel[which], count, host);
}

int biocfd_normdis_download(int id, int which, double *host, int64_t count) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (which < 0 || which >= BN_NNEL || count != b->counts[BC_IB]) { set_error("normdis_download: bad request"); return -1; }
  return download_vec(b->ib.normdis[which], count, host);
}

int biocfd_ghost_download(int id, int which, double *host, int64_t count) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (which < 0 || which >= BGH_N || count != b->counts[BC_TS]) { set_error("ghost_download: bad request"); return -1; }
  return download_vec(b->ib.ghost[which], count, host);
}

int biocfd_elem_download(int id, int which, double *host, int64_t count) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (which < 0 || which >= BE_NELEM || count != b->ib.n_elems) { set_error("elem_download: bad request"); return -1; }
  return download_vec(b->ib.elem[which], count, host);
}

int biocfd_index_ts_download(int id, int64_t *host, int64_t count) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  if (count != b->counts[BC_TS]) { set_error("index_ts_download: bad count"); return -1; }
  return download_vec(b->ib.index_ts, count, host);
}

}  // extern "C"
