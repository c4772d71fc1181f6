// Output sync: the block state the reference's field output writes
// (WriteOutput_corner1.F90:80-112, write_output_ascii; the HDF5 writer
// :21-78 writes the same arrays), assembled on the device and brought to the
// host in one copy.  Per interior cell, i fastest: the cell-centre
// velocities u1 = 0.5 (u(i) + u(i-1)), v1, w1 (the same expressions), p,
// and the tags cell, cell_n, cell_pr.  The host widens the tags to int64.
#include "biocfd_rt.hpp"

namespace bcfd {
namespace {

// stage: 4 doubles per interior cell (u1, v1, w1, p in four planes of n
// entries), then 3 tag bytes per cell packed after them
__global__ void k_output_view(Dims d, const double *__restrict__ u, const double *__restrict__ v,
                              const double *__restrict__ w, const double *__restrict__ p,
                              const uint8_t *__restrict__ cell, const uint8_t *__restrict__ cell_n,
                              const uint8_t *__restrict__ cell_pr, double *__restrict__ stage, int64_t n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
  const int j = blockIdx.y + 2, k = blockIdx.z + 2;
  if (i > d.nx + 1) return;
  const int64_t o = (int64_t)(i - 2) + (int64_t)d.nx * ((int64_t)(j - 2) + (int64_t)d.ny * (int64_t)(k - 2));
  const int64_t c = ix3(d, i, j, k);
  stage[o] = 0.5 * (u[c] + u[ix3(d, i - 1, j, k)]);
  stage[n + o] = 0.5 * (v[c] + v[ix3(d, i, j - 1, k)]);
  stage[2 * n + o] = 0.5 * (w[c] + w[ix3(d, i, j, k - 1)]);
  stage[3 * n + o] = p[c];
  uint8_t *t = reinterpret_cast<uint8_t *>(stage + 4 * n);
  t[o] = cell[c];
  t[n + o] = cell_n[c];
  t[2 * n + o] = cell_pr[ixt(d, i, j, k)];
}

}  // namespace
}  // namespace bcfd

using namespace bcfd;

extern "C" {

int biocfd_output_fields(int id, double *u1, double *v1, double *w1, double *p, int64_t *cell, int64_t *cell_n,
                         int64_t *cell_pr) {
  BCFD_API_LOCK();
  Block *b = get_block(id);
  if (!b) return -1;
  const int64_t n = (int64_t)b->d.nx * b->d.ny * b->d.nz;
  const int64_t words = 4 * n + (3 * n + 7) / 8;   // doubles, then the tag bytes
  if (grow(&b->out_stage, &b->out_cap, words)) return -1;
  hipStream_t s = rt().stream;
  dim3 blk(64, 1, 1), grd(div_up(b->d.nx, 64), b->d.ny, b->d.nz);
  hipLaunchKernelGGL(k_output_view, grd, blk, 0, s, b->d, b->f[BF_U], b->f[BF_V], b->f[BF_W], b->f[BF_P],
                     b->tag[BT_CELL], b->tag[BT_CELL_N], b->tag[BT_CELL_PR], b->out_stage, n);
  BCFD_CHECK_LAUNCH();
  std::vector<double> h((size_t)words);
  BCFD_HIP(hipMemcpyAsync(h.data(), b->out_stage, sizeof(double) * words, hipMemcpyDeviceToHost, s));
  BCFD_HIP(hipStreamSynchronize(s));
  for (int64_t q = 0; q < n; ++q) {
    u1[q] = h[q];
    v1[q] = h[n + q];
    w1[q] = h[2 * n + q];
    p[q] = h[3 * n + q];
  }
  const uint8_t *t = reinterpret_cast<const uint8_t *>(h.data() + 4 * n);
  for (int64_t q = 0; q < n; ++q) {
    cell[q] = t[q];
    cell_n[q] = t[n + q];
    cell_pr[q] = t[2 * n + q];
  }
  return 0;
}

}  // extern "C"
